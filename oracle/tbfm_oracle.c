/* CPU restatement of CTurboBFM's explicit residual path -- TEST INFRASTRUCTURE.
 *
 * The oracle.  Loaded only by tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py, always as the test oracle.  Each function names the reference
 * file:line it restates; operation order follows the reference so that, with
 * -ffp-contract=off on both sides, results agree bit for bit where no libm
 * transcendental is involved.
 *
 * Pinning (what checks each part; the tests are under tests/):
 *   geometry (dual grid, faces, volumes, blockage gradient)     reference build (golden geometry_*),
 *                                                                dual-volume closed form
 *   fluid relations, ideal and table fluid, primitive state     reference build (fluid, lut)
 *   JST / Roe / MUSCL (4 limiters) face fluxes                  reference build (flux_*), Roe
 *                                                                consistency and rotation
 *   boundary fluxes: inlet, outlet + ramp, radial equilibrium,  reference build (boundary,
 *     throttle / mass-flow outlets, inviscid wall               outlets, lut)
 *   transparent end                                             Sod regression solution (0.002 %)
 *   stagnation flux of inlet faces on no-slip walls             isentropic closed form
 *   radial-equilibrium integration                              reference build + free vortex
 *   outlet hub rules (throttle, mass flow), Greitzer plenum     closed forms / known answers;
 *                                                                plenum steps vs reference build
 *   body force: blockage, Hall, lift-drag, Hall-Thollet         reference build per node and over
 *     (off-design, stall), side outputs, time dependence        three serial sweeps (bfm_*)
 *   Gong and 3-D / axisymmetric geometric sources               closed forms (quadratic fields,
 *                                                                wrap planes, solid-body swirl)
 *   viscous gradients, stress tensor, viscosity                 reference build; Poiseuille /
 *   viscous fluxes, no-slip walls, viscous time step            Couette closed forms
 *   time step                                                   closed form on Cartesian grids
 *   face loop, residual assembly, periodic closure              free-stream preservation, closed-
 *                                                                box conservation, node-mate and
 *                                                                sector-rotation properties,
 *                                                                reference regression solutions
 *   RK stage update, norms, convergence test                    recursion closed forms, semantics,
 *                                                                regression runs
 *   turbomachinery figures (mass flow, PRtt, TRtt, ETAtt)       closed forms on uniform planes
 * The reference regression cases match reference.csv within 0.002-0.6 % relative L2,
 * except GER4_FullAnnulus at 1.3-2.7 %, where both codes converge to a non-axisymmetric
 * seam state from axisymmetric inputs; off the hub row and the seam planes it agrees
 * within 0.5-1.1 %.
 */
#define _GNU_SOURCE
#include "tbfm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct { double x, y, z; } v3;
typedef struct { double v[5]; } st5;

static inline v3 mk3(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 add3(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub3(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul3(v3 a, double s) { return mk3(a.x * s, a.y * s, a.z * s); }
static inline v3 div3(v3 a, double s) { return mk3(a.x / s, a.y / s, a.z / s); }
static inline v3 neg3(v3 a) { return mk3(-a.x, -a.y, -a.z); }
static inline double dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double mag3(v3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline v3 cross3(v3 a, v3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* Vector3D::normalized, include/types.hpp:104-107 */
static inline v3 nrm3(v3 a) { double m = mag3(a); return (m == 0) ? mk3(0, 0, 0) : div3(a, m); }

static inline double dmax(double a, double b) { return (a < b) ? b : a; }   /* std::max */
static inline double dmin(double a, double b) { return (b < a) ? b : a; }   /* std::min */

struct tbo_case {
    tbo_params p;
    long N;
    double *x, *y, *z;               /* vertices */
    double *dual;                    /* (NI+1)(NJ+1)(NK+1)*3 */
    double *sI, *sJ, *sK;            /* surfaces */
    double *cI, *cJ, *cK;            /* face centres */
    double *vol;
    double *radius, *theta;
    double *gradb;                   /* blockage gradient 3*N (BFM only) */
    double *blockage, *nax, *nrad, *ntan, *rpm, *present, *nblades, *metal, *slen, *beta0, *kp;
    double *pivot, *stall;                 /* deviationAnglePivot, deviationAngleStall */
    tbo_patch *patches;
    int npatches;
    int *mapI, *mapJ, *mapK;         /* patch index per boundary face, 2 layers each */
    int periodic;
    double per_angle_deg, per_angle_rad;
    double wedge_angle;
    /* outlet pressure profiles (one per radial_equilibrium, throttle or mass_flow patch);
     * prof_radial: integrated radially (else uniform at the hub value) */
    int nprof;
    int prof_patch[8];
    int prof_radial[8];
    /* Greitzer plenum state: pressure, outlet mass flow, temperature (greitzer_model.cpp) */
    double gz_p, gz_mout, gz_T;
    double *prof_p[8], *prof_r[8];
    long prof_n[8];
    int patch_prof[64];
    /* solution */
    double *U0, *U, *R, *dt;
    double *finv, *fvisc, *dev;
    double t_phys;
    /* no-slip walls: node flag and wall velocity of the last no_slip_wall patch holding the
     * node (solver.cpp:369-389); gradients [4 fields][3][N] of a viscous case */
    int* nsw;
    double* wallv;
    double* grad;
    /* look-up-table fluid (FLUID_MODEL = real) */
    tbo_lut lut;
    double dx[TBO_LUT_COUNT], dy[TBO_LUT_COUNT];
    double* lutdata[TBO_LUT_COUNT];
};

#define FI(i, j, k) ((((long)(i) * c->p.nj + (j)) * c->p.nk + (k)) * 3)
#define FJ(i, j, k) ((((long)(i) * (c->p.nj + 1) + (j)) * c->p.nk + (k)) * 3)
#define FK(i, j, k) ((((long)(i) * c->p.nj + (j)) * (c->p.nk + 1) + (k)) * 3)
#define IDX(c, i, j, k) (((long)(i) * (c)->p.nj + (long)(j)) * (c)->p.nk + (long)(k))

static double* dalloc(long n) { return (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

/* ------------------------------------------------------- fluid: look-up table */
/* Table2D::interpolate (include/fluid_real.hpp:34-58): clamp, cell index by truncation,
 * bilinear weights, the four terms summed left to right. */
static double lut_interp(const tbo_case* c, int t, double x, double y) {
    const tbo_table* T = &c->lut.t[t];
    double xc = (x < T->xmin) ? T->xmin : ((T->xmax < x) ? T->xmax : x);   /* std::clamp */
    double yc = (y < T->ymin) ? T->ymin : ((T->ymax < y) ? T->ymax : y);
    double rx = (xc - T->xmin) / c->dx[t];
    double ry = (yc - T->ymin) / c->dy[t];
    size_t i = (size_t)rx, j = (size_t)ry;
    if (i >= (size_t)T->nx - 1) i = (T->nx >= 2) ? (size_t)T->nx - 2 : 0;
    if (j >= (size_t)T->ny - 1) j = (T->ny >= 2) ? (size_t)T->ny - 2 : 0;
    double wx = rx - (double)i, wy = ry - (double)j;
    const double* d = c->lutdata[t];
    double f00 = d[i * T->ny + j], f10 = d[(i + 1) * T->ny + j];
    double f01 = d[i * T->ny + (j + 1)], f11 = d[(i + 1) * T->ny + (j + 1)];
    return (1.0 - wx) * (1.0 - wy) * f00 + wx * (1.0 - wy) * f10 + (1.0 - wx) * wy * f01 + wx * wy * f11;
}
/* FluidReal::computeStaticEnergy_p_rho (src/fluid_real.cpp:227-252): bisection on e */
static double lut_e_p_rho(const tbo_case* c, double p, double rho) {
    double e_low = c->lut.t[TBO_LUT_P].ymin, e_high = c->lut.t[TBO_LUT_P].ymax;
    double p_low = lut_interp(c, TBO_LUT_P, rho, e_low);
    double p_high = lut_interp(c, TBO_LUT_P, rho, e_high);
    if (p <= p_low) return e_low;
    if (p >= p_high) return e_high;
    for (int it = 0; it < 30; ++it) {
        double e_mid = 0.5 * (e_low + e_high);
        double p_mid = lut_interp(c, TBO_LUT_P, rho, e_mid);
        if (fabs(p_mid - p) <= 1e-4 * p) return e_mid;
        if (p_mid < p) e_low = e_mid; else e_high = e_mid;
    }
    return 0.5 * (e_low + e_high);
}
/* FluidBase::computeEnthalpy_p_s (include/fluid_base.hpp) e(p,s) + p/rho(p,s) */
static double lut_h_p_s(const tbo_case* c, double p, double s) {
    return lut_interp(c, TBO_LUT_E_PS, p, s) + p / lut_interp(c, TBO_LUT_RHO_PS, p, s);
}
/* FluidReal::computePressure_rho_T (fluid_real.cpp:264-279) */
static double lut_p_rho_T(const tbo_case* c, double rho, double T) {
    double e_low = c->lut.t[TBO_LUT_T].ymin, e_high = c->lut.t[TBO_LUT_T].ymax;
    for (int it = 0; it < 25; ++it) {
        double e_mid = 0.5 * (e_low + e_high);
        double T_mid = lut_interp(c, TBO_LUT_T, rho, e_mid);
        if (T_mid < T) e_low = e_mid; else e_high = e_mid;
    }
    return lut_interp(c, TBO_LUT_P, rho, 0.5 * (e_low + e_high));
}

double tbo_lut_eval(const tbo_case* c, int which, double x, double y) {
    switch (which) {
        case 0: return lut_interp(c, TBO_LUT_P, x, y);
        case 1: return lut_interp(c, TBO_LUT_T, x, y);
        case 2: return lut_interp(c, TBO_LUT_A, x, y);
        case 3: return lut_interp(c, TBO_LUT_S, x, y);
        case 4: return lut_interp(c, TBO_LUT_DPDRHO, x, y);
        case 5: return lut_interp(c, TBO_LUT_DPDE, x, y);
        case 6: return lut_e_p_rho(c, x, y);
        case 7: return lut_interp(c, TBO_LUT_A, y, lut_e_p_rho(c, x, y));
        case 8: return lut_interp(c, TBO_LUT_RHO_PT, x, y);
        case 9: return lut_interp(c, TBO_LUT_S_PT, x, y);
        case 10: return lut_interp(c, TBO_LUT_RHO_PS, x, y);
        case 11: return lut_interp(c, TBO_LUT_T_PS, x, y);
        case 12: return lut_interp(c, TBO_LUT_E_PS, x, y);
        case 13: return lut_h_p_s(c, x, y);
        case 14: return lut_interp(c, TBO_LUT_E_PT, x, y);
        case 15: return lut_interp(c, TBO_LUT_A_PT, x, y);
        case 16: return lut_p_rho_T(c, x, y);
        default: return NAN;
    }
}

void tbo_set_lut(tbo_case* c, const tbo_lut* lut) {
    c->lut = *lut;
    for (int t = 0; t < TBO_LUT_COUNT; ++t) {
        tbo_table* T = &c->lut.t[t];
        long n = T->nx * T->ny;
        free(c->lutdata[t]);
        c->lutdata[t] = dalloc(n);
        if (T->data && n > 0) memcpy(c->lutdata[t], T->data, sizeof(double) * (size_t)n);
        T->data = c->lutdata[t];
        /* Table2D::init (fluid_real.hpp:22-32) */
        c->dx[t] = (T->nx > 1) ? (T->xmax - T->xmin) / (double)(T->nx - 1) : 1.0;
        c->dy[t] = (T->ny > 1) ? (T->ymax - T->ymin) / (double)(T->ny - 1) : 1.0;
    }
}

/* ------------------------------------------------------------------ fluid */
/* Static energy from velocity and total energy.  Ideal: FluidIdeal override
 * (src/fluid_ideal.cpp:70-77) et - 0.5*pow(|u|,2); real: FluidBase
 * (include/fluid_base.hpp) et - 0.5*u.dot(u). */
static inline double f_e_c(const tbo_case* c, v3 u, double et) {
    if (c->p.fluid_real) return et - 0.5 * dot3(u, u);
    double m = mag3(u);
    return et - 0.5 * (m * m);
}
#define f_e(u, et) f_e_c(c, (u), (et))
/* fluid_ideal.cpp:22-24 / fluid_real.cpp:211-213 */
static inline double f_p_rho_e(const tbo_case* c, double rho, double e) {
    if (c->p.fluid_real) return lut_interp(c, TBO_LUT_P, rho, e);
    return (c->p.gamma - 1) * rho * e;
}
/* fluid_ideal.cpp:90-92 / FluidBase::computePressure_rho_u_et */
static inline double f_p(const tbo_case* c, double rho, v3 u, double et) { return f_p_rho_e(c, rho, f_e(u, et)); }
/* computeSoundSpeed_p_rho: ideal sqrt(gamma p / rho) (fluid_ideal.cpp:70-72); real
 * a(rho, e(p, rho)) with the bisection for e (fluid_real.cpp:254-257) */
static inline double f_a_p_rho(const tbo_case* c, double p, double rho) {
    if (c->p.fluid_real) return lut_interp(c, TBO_LUT_A, rho, lut_e_p_rho(c, p, rho));
    return sqrt(c->p.gamma * p / rho);
}
/* computeSoundSpeed_rho_u_et: ideal override via p; real a(rho, e) */
static inline double f_a(const tbo_case* c, double rho, v3 u, double et) {
    if (c->p.fluid_real) return lut_interp(c, TBO_LUT_A, rho, f_e(u, et));
    double p = f_p(c, rho, u, et);
    return f_a_p_rho(c, p, rho);
}
/* computeSoundSpeed_rho_e: ideal sqrt(gamma (gamma-1) e); real a(rho, e) */
static inline double f_a_rho_e(const tbo_case* c, double rho, double e) {
    if (c->p.fluid_real) return lut_interp(c, TBO_LUT_A, rho, e);
    return sqrt(c->p.gamma * (c->p.gamma - 1.0) * e);
}
/* computeDpDrho_e / computeDpDe_rho: ideal (gamma-1) e, (gamma-1) rho; real tables */
static inline double f_chi(const tbo_case* c, double rho, double e) {
    if (c->p.fluid_real) return lut_interp(c, TBO_LUT_DPDRHO, rho, e);
    return (c->p.gamma - 1.0) * e;
}
static inline double f_kappa(const tbo_case* c, double rho, double e) {
    if (c->p.fluid_real) return lut_interp(c, TBO_LUT_DPDE, rho, e);
    return (c->p.gamma - 1.0) * rho;
}
/* fluid_ideal.cpp:111-115 / FluidBase::computeTotalEnthalpy_rho_u_et: et + p/rho */
static inline double f_ht(const tbo_case* c, double rho, v3 u, double et) {
    double e = f_e(u, et);
    double p = f_p_rho_e(c, rho, e);
    return et + p / rho;
}
static inline double f_T(const tbo_case* c, double rho, v3 u, double et) {
    if (c->p.fluid_real) return lut_interp(c, TBO_LUT_T, rho, f_e(u, et));
    return f_p(c, rho, u, et) / (rho * c->p.r_gas);
}
static inline double f_cp(const tbo_case* c) { return (c->p.gamma * c->p.r_gas) / (c->p.gamma - 1); }
/* total pressure: ideal isentropic relation; real FluidReal::computeTotalPressure_rho_u_et
 * (fluid_real.cpp:311-334): bisection on p_t >= p with h(p_t, s) = h_t */
static double f_pt(const tbo_case* c, double rho, v3 u, double et) {
    if (c->p.fluid_real) {
        double e = f_e(u, et);
        double p = lut_interp(c, TBO_LUT_P, rho, e);
        double s = lut_interp(c, TBO_LUT_S, rho, e);
        double ht = et + p / rho;
        double p_low = p, p_high = c->lut.t[TBO_LUT_RHO_PS].xmax;
        for (int it = 0; it < 25; ++it) {
            double p_mid = 0.5 * (p_low + p_high);
            double rho_mid = lut_interp(c, TBO_LUT_RHO_PS, p_mid, s);
            double e_mid = lut_interp(c, TBO_LUT_E_PS, p_mid, s);
            double h_mid = e_mid + p_mid / rho_mid;
            if (h_mid < ht) p_low = p_mid; else p_high = p_mid;
        }
        return 0.5 * (p_low + p_high);
    }
    double g = c->p.gamma;
    double p = f_p(c, rho, u, et);
    double M = mag3(u) / f_a(c, rho, u, et);
    return p * pow(1.0 + (g - 1.0) / 2.0 * M * M, g / (g - 1));
}
/* real: FluidReal::computeTotalTemperature_rho_u_et (fluid_real.cpp:336-341) */
static double f_Tt(const tbo_case* c, double rho, v3 u, double et) {
    if (c->p.fluid_real) {
        double pt = f_pt(c, rho, u, et);
        double e = f_e(u, et);
        double s = lut_interp(c, TBO_LUT_S, rho, e);
        return lut_interp(c, TBO_LUT_T_PS, pt, s);
    }
    double g = c->p.gamma;
    double T = f_T(c, rho, u, et);
    double M = mag3(u) / f_a(c, rho, u, et);
    return T * (1.0 + (g - 1.0) / 2.0 * M * M);
}
static double f_s(const tbo_case* c, double rho, v3 u, double et) {
    if (c->p.fluid_real) return lut_interp(c, TBO_LUT_S, rho, f_e(u, et));
    double p = f_p(c, rho, u, et);
    double T = f_T(c, rho, u, et);
    return f_cp(c) * log(T / 288.15) - c->p.r_gas * log(p / 101325);
}

void tbo_fluid_state(const tbo_case* c, const double* w, double* out) {
    v3 u = mk3(w[1], w[2], w[3]);
    out[0] = f_p(c, w[0], u, w[4]);
    out[1] = f_a(c, w[0], u, w[4]);
    out[2] = f_ht(c, w[0], u, w[4]);
    out[3] = f_T(c, w[0], u, w[4]);
    out[4] = f_pt(c, w[0], u, w[4]);
    out[5] = f_Tt(c, w[0], u, w[4]);
    out[6] = f_s(c, w[0], u, w[4]);
}

/* ------------------------------------------------------------------ math utils */
/* src/math_utils.cpp:49-57 */
static inline st5 prim(st5 U) {
    st5 W;
    W.v[0] = U.v[0];
    W.v[1] = U.v[1] / U.v[0];
    W.v[2] = U.v[2] / U.v[0];
    W.v[3] = U.v[3] / U.v[0];
    W.v[4] = U.v[4] / U.v[0];
    return W;
}
/* src/math_utils.cpp:71-86 */
static inline st5 phys_flux(const tbo_case* c, st5 W, v3 S) {
    v3 n = div3(S, mag3(S));
    v3 u = mk3(W.v[1], W.v[2], W.v[3]);
    double un = dot3(u, n);
    double p = f_p(c, W.v[0], u, W.v[4]);
    double ht = f_ht(c, W.v[0], u, W.v[4]);
    st5 F;
    F.v[0] = W.v[0] * un;
    F.v[1] = W.v[0] * un * u.x + p * n.x;
    F.v[2] = W.v[0] * un * u.y + p * n.y;
    F.v[3] = W.v[0] * un * u.z + p * n.z;
    F.v[4] = W.v[0] * un * ht;
    return F;
}
/* math_utils.cpp:109-122 */
static inline st5 rotx(st5 s, double cs, double sn) {
    st5 r = s;
    r.v[2] = s.v[2] * cs - s.v[3] * sn;
    r.v[3] = s.v[2] * sn + s.v[3] * cs;
    return r;
}
/* math_utils.cpp:124-138 */
static inline v3 cyl_from_cart(v3 v, double cs, double sn) {
    return mk3(v.x, v.y * cs + v.z * sn, -v.y * sn + v.z * cs);
}
static inline v3 cart_from_cyl(v3 v, double cs, double sn) {
    return mk3(v.x, v.y * cs - v.z * sn, v.y * sn + v.z * cs);
}
/* math_utils.cpp:141-170 */
void tbo_radial_equilibrium(const double* rho, const double* ut, const double* r, int n,
                            double hub, double* p) {
    p[0] = hub;
    for (int i = 1; i < n; ++i) {
        double dr = r[i] - r[i - 1];
        double d1 = rho[i - 1] * ut[i - 1] * ut[i - 1] / r[i - 1];
        double d2 = rho[i] * ut[i] * ut[i] / r[i];
        p[i] = p[i - 1] + 0.5 * (d1 + d2) * dr;
    }
}

static inline st5 ld(const tbo_case* c, const double* U, long idx) {
    st5 s;
    long N = c->N;
    for (int v = 0; v < 5; ++v) s.v[v] = U[v * N + idx];
    return s;
}

/* ------------------------------------------------------------------ face fluxes */
/* src/advection_jst.cpp:40-57 */
static inline double jst_r(const tbo_case* c, st5 W) {
    v3 u = mk3(W.v[1], W.v[2], W.v[3]);
    double vm = mag3(u);
    double a = f_a(c, W.v[0], u, W.v[4]);
    return vm + a;
}
static inline double jst_s(const tbo_case* c, st5 W1, st5 W2, st5 W3) {
    double p1 = f_p(c, W1.v[0], mk3(W1.v[1], W1.v[2], W1.v[3]), W1.v[4]);
    double p2 = f_p(c, W2.v[0], mk3(W2.v[1], W2.v[2], W2.v[3]), W2.v[4]);
    double p3 = f_p(c, W3.v[0], mk3(W3.v[1], W3.v[2], W3.v[3]), W3.v[4]);
    return fabs((p1 - 2 * p2 + p3) / (p1 + 2 * p2 + p3));
}
/* src/advection_jst.cpp:4-38, constants include/advection_jst.hpp:23 */
static st5 jst_flux(const tbo_case* c, st5 Ull, st5 Ul, st5 Ur, st5 Urr, v3 S) {
    const double k2 = 0.5, k4 = 0.02, c4 = 2.0;
    st5 Wll = prim(Ull), Wl = prim(Ul), Wr = prim(Ur), Wrr = prim(Urr);
    st5 Wavg;
    for (int v = 0; v < 5; ++v) Wavg.v[v] = (Wl.v[v] + Wr.v[v]) * 0.5;
    double r0 = jst_r(c, Wl), r1 = jst_r(c, Wr);
    double s0 = jst_s(c, Wll, Wl, Wr), s1 = jst_s(c, Wl, Wr, Wrr);
    double r = dmax(r0, r1);
    double s = dmax(s0, s1);
    double psi2 = k2 * s * r;
    double psi4 = dmax(0.0, k4 * r - c4 * psi2);
    st5 F = phys_flux(c, Wavg, S);
    for (int v = 0; v < 5; ++v) {
        double d1 = Ur.v[v] - Ul.v[v];
        double d3 = ((Urr.v[v] - Ur.v[v]) - d1 * 2) + (Ul.v[v] - Ull.v[v]);
        double dis = d1 * psi2 - d3 * psi4;
        F.v[v] -= dis;
    }
    return F;
}

/* src/advection_base.cpp:31-65 */
static inline double smooth1(double l, double cc, double r) {
    double dR = r - cc;
    double dL = cc - l;
    return (fabs(dR) < 1E-12) ? 0.0 : dL / dR;
}
static inline double limiter1(int lim, double s) {
    switch (lim) {
        case TBO_LIM_VAN_ALBADA: return (s <= 0.0) ? 0.0 : (s * s + s) / (1.0 + s * s);
        case TBO_LIM_VAN_LEER: return (s + fabs(s)) / (1.0 + fabs(s));
        case TBO_LIM_MIN_MOD: return dmax(0.0, dmin(1.0, s));
        default: return 0.0;
    }
}
/* src/advection_base.cpp:10-29 */
static void muscl(int lim, st5* Wll, st5* Wl, st5* Wr, st5* Wrr) {
    double sl[5], sr[5];
    for (int v = 0; v < 5; ++v) {
        sl[v] = smooth1(Wll->v[v], Wl->v[v], Wr->v[v]);
        sr[v] = smooth1(Wl->v[v], Wr->v[v], Wrr->v[v]);
    }
    for (int v = 0; v < 5; ++v) {
        double ll = limiter1(lim, sl[v]);
        double lr = limiter1(lim, sr[v]);
        double wl = Wl->v[v], wr = Wr->v[v], wrr = Wrr->v[v];
        Wl->v[v] = wl + ll * (wr - wl) * 0.5;
        Wr->v[v] = wr - lr * (wrr - wr) * 0.5;
    }
}

/* src/advection_roe.cpp:146-149 */
static inline double roe_avg(double rl, double rr, double pl, double pr) {
    return (sqrt(rl) * pl + sqrt(rr) * pr) / (sqrt(rl) + sqrt(rr));
}
/* src/advection_roe.cpp:6-221, entropy fix 0.01 include/advection_roe.hpp:63 */
static st5 roe_flux(const tbo_case* c, st5 Ull, st5 Ul, st5 Ur, st5 Urr, v3 S) {
    const double g = c->p.gamma;
    st5 Wl = prim(Ul), Wr = prim(Ur), Wrr = prim(Urr), Wll = prim(Ull);
    if (c->p.muscl) muscl(c->p.limiter, &Wll, &Wl, &Wr, &Wrr);
    /* computeNormalTriad (:41-58) */
    v3 n1 = nrm3(S), n2;
    if (n1.x == 0 && n1.y == 0) n2 = mk3(1, 0, 0);
    else n2 = mk3(-n1.y, n1.x, 0);
    n2 = nrm3(n2);
    v3 n3 = cross3(n1, n2);
    /* computeRotatedPrimitive (:60-64) */
    st5 L, R;
    v3 ul = mk3(Wl.v[1], Wl.v[2], Wl.v[3]), ur = mk3(Wr.v[1], Wr.v[2], Wr.v[3]);
    L.v[0] = Wl.v[0]; L.v[1] = dot3(ul, n1); L.v[2] = dot3(ul, n2); L.v[3] = dot3(ul, n3); L.v[4] = Wl.v[4];
    R.v[0] = Wr.v[0]; R.v[1] = dot3(ur, n1); R.v[2] = dot3(ur, n2); R.v[3] = dot3(ur, n3); R.v[4] = Wr.v[4];
    /* computeRoeAvgVariables (:66-144) */
    double rhoL = L.v[0], rhoR = R.v[0];
    double u1L = L.v[1], u1R = R.v[1], u2L = L.v[2], u2R = R.v[2], u3L = L.v[3], u3R = R.v[3];
    double etL = L.v[4], etR = R.v[4];
    double htL = f_ht(c, rhoL, mk3(u1L, u2L, u3L), etL);
    double htR = f_ht(c, rhoR, mk3(u1R, u2R, u3R), etR);
    double rhoA = sqrt(rhoL * rhoR);
    double u1A = roe_avg(rhoL, rhoR, u1L, u1R);
    double u2A = roe_avg(rhoL, rhoR, u2L, u2R);
    double u3A = roe_avg(rhoL, rhoR, u3L, u3R);
    double htA = roe_avg(rhoL, rhoR, htL, htR);
    double eL = etL - 0.5 * (u1L * u1L + u2L * u2L + u3L * u3L);
    double eR = etR - 0.5 * (u1R * u1R + u2R * u2R + u3R * u3R);
    double pL = f_p_rho_e(c, rhoL, eL), pR = f_p_rho_e(c, rhoR, eR);
    double chiL = f_chi(c, rhoL, eL), kapL = f_kappa(c, rhoL, eL);
    double chiR = f_chi(c, rhoR, eR), kapR = f_kappa(c, rhoR, eR);
    double barChi = 0.5 * (chiL + chiR), barKap = 0.5 * (kapL + kapR);
    double barRho = 0.5 * (rhoL + rhoR), barE = 0.5 * (eL + eR);
    double dP = pR - pL, dRho = rhoR - rhoL, dE = eR - eL;
    double dRhoHat = (barRho > 1e-12) ? dRho / barRho : 0.0;
    double dEHat = (fabs(barE) > 1e-12) ? dE / fabs(barE) : 0.0;
    double dNorm2 = dRhoHat * dRhoHat + dEHat * dEHat;
    double chiT = barChi, kapT = barKap;
    if (dNorm2 > 1e-12) {
        double res = dP - (barChi * dRho + barKap * dE);
        chiT += (res * dRhoHat) / (dNorm2 * barRho);
        kapT += (res * dEHat) / (dNorm2 * fabs(barE));
    }
    if (barKap > 0.0 && kapT <= 0.0) kapT = 0.01 * barKap;
    double kapP = (rhoA > 1e-12) ? kapT / rhoA : 0.4;
    double eA = roe_avg(rhoL, rhoR, eL, eR);
    double chiP = chiT - kapP * eA;
    double q2 = u1A * u1A + u2A * u2A + u3A * u3A;
    double hA = htA - 0.5 * q2;
    double a2 = chiP + kapP * hA;
    if (a2 <= 0.0) {
        double aL = f_a_rho_e(c, rhoL, eL), aR = f_a_rho_e(c, rhoR, eR);
        a2 = 0.5 * (aL * aL + aR * aR);
    }
    double aA = sqrt(a2);
    /* computeEigenvalues / computeEigenvectors (:151-167) */
    double lam[5] = {u1A - aA, u1A, u1A, u1A, u1A + aA};
    double e5 = htA - (aA * aA) / kapP;
    double rv[5][5] = {{1.0, u1A - aA, u2A, u3A, htA - aA * u1A},
                       {1.0, u1A, u2A, u3A, e5},
                       {0.0, 0.0, 1.0, 0.0, u2A},
                       {0.0, 0.0, 0.0, 1.0, u3A},
                       {1.0, u1A + aA, u2A, u3A, htA + aA * u1A}};
    /* computeWaveStrengths (:169-182) */
    double dRho2 = R.v[0] - L.v[0];
    double dU1 = R.v[1] - L.v[1], dU2 = R.v[2] - L.v[2], dU3 = R.v[3] - L.v[3];
    double pl2 = f_p(c, L.v[0], mk3(L.v[1], L.v[2], L.v[3]), L.v[4]);
    double pr2 = f_p(c, R.v[0], mk3(R.v[1], R.v[2], R.v[3]), R.v[4]);
    double dP2 = pr2 - pl2;
    double w[5];
    w[0] = 1.0 / (2.0 * aA * aA) * (dP2 - rhoA * aA * dU1);
    w[1] = dRho2 - dP2 / (aA * aA);
    w[2] = rhoA * dU2;
    w[3] = rhoA * dU3;
    w[4] = 1.0 / (2.0 * aA * aA) * (dP2 + rhoA * aA * dU1);
    /* assembleTotalFlux (:184-221) */
    v3 ex = mk3(1.0, 0.0, 0.0);
    st5 FL = phys_flux(c, L, ex), FR = phys_flux(c, R, ex);
    double fe[5];
    for (int i = 0; i < 5; ++i) fe[i] = (fabs(lam[i]) < 0.01) ? 0.01 : fabs(lam[i]);
    st5 Fr;
    for (int i = 0; i < 5; ++i) Fr.v[i] = (FL.v[i] + FR.v[i]) * 0.5;
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) Fr.v[i] -= 0.5 * w[j] * fe[j] * rv[j][i];
    v3 xv = mk3(1.0, 0.0, 0.0), yv = mk3(0.0, 1.0, 0.0), zv = mk3(0.0, 0.0, 1.0);
    st5 F;
    F.v[0] = Fr.v[0];
    F.v[1] = Fr.v[1] * dot3(n1, xv) + Fr.v[2] * dot3(n2, xv) + Fr.v[3] * dot3(n3, xv);
    F.v[2] = Fr.v[1] * dot3(n1, yv) + Fr.v[2] * dot3(n2, yv) + Fr.v[3] * dot3(n3, yv);
    F.v[3] = Fr.v[1] * dot3(n1, zv) + Fr.v[2] * dot3(n2, zv) + Fr.v[3] * dot3(n3, zv);
    F.v[4] = Fr.v[4];
    return F;
}

static inline st5 face_flux(const tbo_case* c, st5 Ull, st5 Ul, st5 Ur, st5 Urr, v3 S) {
    return (c->p.scheme == TBO_ROE) ? roe_flux(c, Ull, Ul, Ur, Urr, S) : jst_flux(c, Ull, Ul, Ur, Urr, S);
}

void tbo_face_flux(const tbo_case* c, const double* a, const double* b, const double* d,
                   const double* e, const double* S, double* out) {
    st5 A, B, D, E;
    memcpy(A.v, a, 40); memcpy(B.v, b, 40); memcpy(D.v, d, 40); memcpy(E.v, e, 40);
    st5 F = face_flux(c, A, B, D, E, mk3(S[0], S[1], S[2]));
    memcpy(out, F.v, 40);
}

/* ------------------------------------------------------------ boundary fluxes */
/* src/boundary_base.cpp:7-124 (ideal-gas branch) */
static st5 inlet_subsonic(const tbo_case* c, st5 Uint, v3 S, double pt, double Tt, v3 dir) {
    const double g = c->p.gamma, Rg = c->p.r_gas;
    st5 W = prim(Uint);
    v3 u = mk3(W.v[1], W.v[2], W.v[3]);
    double aInt = f_a(c, W.v[0], u, W.v[4]);
    v3 n = div3(S, mag3(S));
    double unInt = dot3(u, n);
    double Z = W.v[0] * aInt;
    double pInt = f_p(c, W.v[0], u, W.v[4]);
    double Cg = dot3(dir, n);
    if (Cg >= 0.0) Cg = -1.0;
    double Rout = pInt + Z * unInt;
    if (c->p.fluid_real) {
        /* boundary_base.cpp:77-111: entropy of the inlet total state, enthalpy search */
        double sIn = lut_interp(c, TBO_LUT_S_PT, pt, Tt);
        double htIn = lut_h_p_s(c, pt, sIn);
        double pLow = 0.05 * pt, pHigh = 0.999999 * pt;
        double pB = 0.5 * (pLow + pHigh);
        for (int it = 0; it < 30; ++it) {
            double pMid = 0.5 * (pLow + pHigh);
            double hMid = lut_h_p_s(c, pMid, sIn);
            double VMid = sqrt(2.0 * dmax(0.0, htIn - hMid));
            double res = pMid + Z * Cg * VMid - Rout;
            if (fabs(res) < 1e-4 * pt) { pB = pMid; break; }
            if (res < 0.0) pLow = pMid; else pHigh = pMid;
            pB = pMid;
        }
        double hB = lut_h_p_s(c, pB, sIn);
        double VB = sqrt(2.0 * dmax(0.0, htIn - hB));
        v3 vB = mul3(dir, VB);
        double rhoB = lut_interp(c, TBO_LUT_RHO_PS, pB, sIn);
        double eB = lut_interp(c, TBO_LUT_E_PS, pB, sIn);
        double etB = eB + 0.5 * dot3(vB, vB);
        st5 Wb = {{rhoB, vB.x, vB.y, vB.z, etB}};
        return phys_flux(c, Wb, S);
    }
    double cp = g * Rg / (g - 1.0);
    double htIn = cp * Tt;
    double pLow = 0.05 * pt, pHigh = 0.999999 * pt;
    double pB = 0.5 * (pLow + pHigh);
    for (int it = 0; it < 30; ++it) {
        double pMid = 0.5 * (pLow + pHigh);
        double TMid = Tt * pow(pMid / pt, (g - 1.0) / g);
        double hMid = cp * TMid;
        double VMid = sqrt(2.0 * dmax(0.0, htIn - hMid));
        double res = pMid + Z * Cg * VMid - Rout;
        if (fabs(res) < 1e-4 * pt) { pB = pMid; break; }
        if (res < 0.0) pLow = pMid; else pHigh = pMid;
        pB = pMid;
    }
    double TB = Tt * pow(pB / pt, (g - 1.0) / g);
    double hB = cp * TB;
    double VB = sqrt(2.0 * dmax(0.0, htIn - hB));
    v3 vB = mul3(dir, VB);
    double rhoB = pB / (Rg * TB);
    double eB = pB / ((g - 1) * rhoB);
    double etB = eB + 0.5 * dot3(vB, vB);
    st5 Wb = {{rhoB, vB.x, vB.y, vB.z, etB}};
    return phys_flux(c, Wb, S);
}
/* src/config.cpp:430-439 */
static double ramped_pressure(const tbo_case* c, long it, double pout) {
    double ratio = (double)it / c->p.ramp_iterations;
    if (ratio >= 1.0) return pout;
    return c->p.init_pressure + (pout - c->p.init_pressure) * ratio;
}
/* src/boundary_base.cpp:181-224 (ideal-gas branch) */
static st5 outlet_flux(const tbo_case* c, st5 Uint, v3 S, double pb, long it) {
    st5 W = prim(Uint);
    v3 u = mk3(W.v[1], W.v[2], W.v[3]);
    double rho = W.v[0];
    double p = f_p(c, rho, u, W.v[4]);
    double a = f_a_p_rho(c, p, W.v[0]);
    if (mag3(u) >= a) return phys_flux(c, W, S);
    double pB = ramped_pressure(c, it, pb);
    double rhoB, eB;
    if (c->p.fluid_real) {
        /* boundary_base.cpp:205-211: isentropic extrapolation along the interior entropy */
        double e_int = W.v[4] - 0.5 * dot3(u, u);
        double s_int = lut_interp(c, TBO_LUT_S, rho, e_int);
        rhoB = lut_interp(c, TBO_LUT_RHO_PS, pB, s_int);
        eB = lut_interp(c, TBO_LUT_E_PS, pB, s_int);
    } else {
        rhoB = pB * rho / p;
        eB = pB / ((c->p.gamma - 1) * rhoB);
    }
    double etB = eB + 0.5 * dot3(u, u);
    st5 Wb = {{rhoB, u.x, u.y, u.z, etB}};
    return phys_flux(c, Wb, S);
}
/* src/boundary_inviscid_wall.cpp:5-30 */
static st5 wall_flux(const tbo_case* c, st5 Uint, v3 S) {
    st5 W = prim(Uint);
    v3 u = mk3(W.v[1], W.v[2], W.v[3]);
    double p = f_p(c, W.v[0], u, W.v[4]);
    v3 n = div3(S, mag3(S));
    double un = dot3(u, n);
    double a = f_a_p_rho(c, p, W.v[0]);
    double pw = p + W.v[0] * un * a;
    if (pw < 0.01 * p) pw = 0.01 * p;
    st5 F = {{0.0, pw * n.x, pw * n.y, pw * n.z, 0.0}};
    return F;
}
/* BoundaryBase::computeStagnationWallFlux (boundary_base.cpp:126-180), ideal gas: the
 * wall's stagnation pressure on the face, no mass or energy flux */
static st5 stagnation_wall_flux(const tbo_case* c, v3 S, double pt, double Tt, v3 uw) {
    double um = mag3(uw), pb = pt;
    if (um > 0.0) {
        double g = c->p.gamma, Rg = c->p.r_gas;
        double cp = g * Rg / (g - 1.0);
        double Tb = Tt - 0.5 * um * um / cp;
        if (Tb < 0.01 * Tt) Tb = 0.01 * Tt;
        pb = pt * pow(Tb / Tt, g / (g - 1.0));
    }
    v3 n = div3(S, mag3(S));
    st5 F = {{0.0, pb * n.x, pb * n.y, pb * n.z, 0.0}};
    return F;
}
/* src/boundary_inlet.cpp:4-43.  The face indices name the node checked against the no-slip
 * walls (boundary_base.hpp:44-56): a start face is its node, an end face lies outside the
 * node range in its direction and never matches. */
static st5 inlet_flux(const tbo_case* c, const double* vals, st5 Uint, v3 S, long fi, long fj, long fk) {
    if (c->nsw && fi < c->p.ni && fj < c->p.nj && fk < c->p.nk) {
        long w = IDX(c, fi, fj, fk);
        if (c->nsw[w]) return stagnation_wall_flux(c, S, vals[0], vals[1], mk3(c->wallv[3 * w], c->wallv[3 * w + 1], c->wallv[3 * w + 2]));
    }
    v3 dir = mk3(vals[2], vals[3], vals[4]);
    if (c->p.inlet_cylindrical) {
        long idx = IDX(c, fi, fj, fk);
        double th = c->theta[idx];
        dir = cart_from_cyl(dir, cos(th), sin(th));
    }
    if (dir.x == 1.0 && dir.y == 1.0 && dir.z == 1.0) dir = neg3(S);
    dir = div3(dir, mag3(dir));
    return inlet_subsonic(c, Uint, S, vals[0], vals[1], dir);
}

static st5 bc_flux(const tbo_case* c, int pidx, st5 Uint, v3 S, long fi, long fj, long fk, long it) {
    const tbo_patch* pa = &c->patches[pidx];
    st5 z = {{0, 0, 0, 0, 0}};
    switch (pa->type) {
        case TBO_BC_INLET: return inlet_flux(c, pa->values, Uint, S, fi, fj, fk);
        case TBO_BC_OUTLET: return outlet_flux(c, Uint, S, pa->values[0], it);
        case TBO_BC_RADIAL_EQ:
        case TBO_BC_THROTTLE: {
            /* boundary_outlet_radial_equilibrium.cpp:5-16, boundary_outlet_throttle.cpp:5-16:
             * profile[j of the face] */
            const double* prof = c->prof_p[c->patch_prof[pidx]];
            return outlet_flux(c, Uint, S, prof[fj], it);
        }
        case TBO_BC_MASS_FLOW: {
            /* boundary_outlet_mass_flow.cpp:4-19: j of the face, clamped to the profile */
            int q = c->patch_prof[pidx];
            long idx = fj;
            if (idx >= c->prof_n[q]) idx = c->prof_n[q] - 1;
            return outlet_flux(c, Uint, S, c->prof_p[q][idx], it);
        }
        case TBO_BC_WALL:
        case TBO_BC_NO_SLIP: return wall_flux(c, Uint, S);   /* solver.cpp:391-399 */
        /* boundary_transparent.cpp:5-16: the scheme's flux with the internal state on all
         * four positions */
        case TBO_BC_TRANSPARENT: return face_flux(c, Uint, Uint, Uint, Uint, S);
        default: return z;
    }
}

void tbo_boundary_flux(const tbo_case* c, int kind, const double* values, const double* Ui,
                       const double* S, const long* face, long iter, const double* profile,
                       double* out) {
    st5 U;
    memcpy(U.v, Ui, 40);
    v3 s = mk3(S[0], S[1], S[2]);
    st5 F;
    if (kind == 0) F = inlet_flux(c, values, U, s, face[0], face[1], face[2]);
    else if (kind == 1) F = outlet_flux(c, U, s, values[0], iter);
    else if (kind == 2) F = outlet_flux(c, U, s, profile[face[1]], iter);
    else F = wall_flux(c, U, s);
    memcpy(out, F.v, 40);
}

/* ------------------------------------------------------------------- geometry */
static void surf_center(v3 p1, v3 p2, v3 p3, v3 p4, double* S, double* C) {
    /* src/mesh.cpp:165-190 */
    v3 a1 = sub3(p2, p1), b1 = sub3(p3, p2), a2 = sub3(p3, p1), b2 = sub3(p4, p3);
    v3 s1 = div3(cross3(a1, b1), 2.0), s2 = div3(cross3(a2, b2), 2.0);
    v3 n = add3(s1, s2);
    v3 c1 = div3(add3(add3(p1, p2), p3), 3.0);
    v3 c2 = div3(add3(add3(p1, p3), p4), 3.0);
    v3 ce = div3(add3(mul3(c1, mag3(s1)), mul3(c2, mag3(s2))), mag3(s1) + mag3(s2));
    S[0] = n.x; S[1] = n.y; S[2] = n.z;
    C[0] = ce.x; C[1] = ce.y; C[2] = ce.z;
}

static inline v3 vtx(const tbo_case* c, long i, long j, long k) {
    long id = IDX(c, i, j, k);
    return mk3(c->x[id], c->y[id], c->z[id]);
}

static void build_geometry(tbo_case* c) {
    const long NI = c->p.ni, NJ = c->p.nj, NK = c->p.nk;
    const long DI = NI + 1, DJ = NJ + 1, DK = NK + 1;
    c->dual = dalloc(DI * DJ * DK * 3);
#define DN(i, j, k) (&c->dual[(((long)(i) * DJ + (j)) * DK + (k)) * 3])
#define SETD(i, j, k, v) do { double* _d = DN(i, j, k); _d[0] = (v).x; _d[1] = (v).y; _d[2] = (v).z; } while (0)
    if (c->p.topology == TBO_TOPO_3D) {
        /* src/mesh.cpp:478-557 */
        for (long i = 1; i < DI - 1; i++)
            for (long j = 1; j < DJ - 1; j++)
                for (long k = 1; k < DK - 1; k++) {
                    v3 s = add3(add3(add3(add3(add3(add3(add3(vtx(c, i - 1, j - 1, k - 1), vtx(c, i, j - 1, k - 1)),
                        vtx(c, i - 1, j, k - 1)), vtx(c, i - 1, j - 1, k)), vtx(c, i, j, k - 1)), vtx(c, i, j - 1, k)),
                        vtx(c, i - 1, j, k)), vtx(c, i, j, k));
                    SETD(i, j, k, div3(s, 8.0));
                }
        SETD(0, 0, 0, vtx(c, 0, 0, 0));
        SETD(DI - 1, 0, 0, vtx(c, NI - 1, 0, 0));
        SETD(0, DJ - 1, 0, vtx(c, 0, NJ - 1, 0));
        SETD(0, 0, DK - 1, vtx(c, 0, 0, NK - 1));
        SETD(DI - 1, DJ - 1, 0, vtx(c, NI - 1, NJ - 1, 0));
        SETD(DI - 1, 0, DK - 1, vtx(c, NI - 1, 0, NK - 1));
        SETD(0, DJ - 1, DK - 1, vtx(c, 0, NJ - 1, NK - 1));
        SETD(DI - 1, DJ - 1, DK - 1, vtx(c, NI - 1, NJ - 1, NK - 1));
        for (long i = 1; i < DI - 1; i++) {
            SETD(i, 0, 0, div3(add3(vtx(c, i - 1, 0, 0), vtx(c, i, 0, 0)), 2.0));
            SETD(i, DJ - 1, 0, div3(add3(vtx(c, i - 1, NJ - 1, 0), vtx(c, i, NJ - 1, 0)), 2.0));
            SETD(i, 0, DK - 1, div3(add3(vtx(c, i - 1, 0, NK - 1), vtx(c, i, 0, NK - 1)), 2.0));
            SETD(i, DJ - 1, DK - 1, div3(add3(vtx(c, i - 1, NJ - 1, NK - 1), vtx(c, i, NJ - 1, NK - 1)), 2.0));
        }
        for (long j = 1; j < DJ - 1; j++) {
            SETD(0, j, 0, div3(add3(vtx(c, 0, j - 1, 0), vtx(c, 0, j, 0)), 2.0));
            SETD(DI - 1, j, 0, div3(add3(vtx(c, NI - 1, j - 1, 0), vtx(c, NI - 1, j, 0)), 2.0));
            SETD(0, j, DK - 1, div3(add3(vtx(c, 0, j - 1, NK - 1), vtx(c, 0, j, NK - 1)), 2.0));
            SETD(DI - 1, j, DK - 1, div3(add3(vtx(c, NI - 1, j - 1, NK - 1), vtx(c, NI - 1, j, NK - 1)), 2.0));
        }
        for (long k = 1; k < DK - 1; k++) {
            SETD(0, 0, k, div3(add3(vtx(c, 0, 0, k - 1), vtx(c, 0, 0, k)), 2.0));
            SETD(DI - 1, 0, k, div3(add3(vtx(c, NI - 1, 0, k - 1), vtx(c, NI - 1, 0, k)), 2.0));
            SETD(0, DJ - 1, k, div3(add3(vtx(c, 0, NJ - 1, k - 1), vtx(c, 0, NJ - 1, k)), 2.0));
            SETD(DI - 1, DJ - 1, k, div3(add3(vtx(c, NI - 1, NJ - 1, k - 1), vtx(c, NI - 1, NJ - 1, k)), 2.0));
        }
        for (long j = 1; j < DJ - 1; j++)
            for (long k = 1; k < DK - 1; k++) {
                SETD(0, j, k, div3(add3(add3(add3(vtx(c, 0, j - 1, k - 1), vtx(c, 0, j, k - 1)), vtx(c, 0, j - 1, k)), vtx(c, 0, j, k)), 4.0));
                SETD(DI - 1, j, k, div3(add3(add3(add3(vtx(c, NI - 1, j - 1, k - 1), vtx(c, NI - 1, j, k - 1)),
                     vtx(c, NI - 1, j - 1, k)), vtx(c, NI - 1, j, k)), 4.0));
            }
        for (long i = 1; i < DI - 1; i++)
            for (long k = 1; k < DK - 1; k++) {
                SETD(i, 0, k, div3(add3(add3(add3(vtx(c, i - 1, 0, k - 1), vtx(c, i, 0, k - 1)), vtx(c, i - 1, 0, k)), vtx(c, i, 0, k)), 4.0));
                SETD(i, DJ - 1, k, div3(add3(add3(add3(vtx(c, i - 1, NJ - 1, k - 1), vtx(c, i, NJ - 1, k - 1)),
                     vtx(c, i - 1, NJ - 1, k)), vtx(c, i, NJ - 1, k)), 4.0));
            }
        for (long i = 1; i < DI - 1; i++)
            for (long j = 1; j < DJ - 1; j++) {
                SETD(i, j, 0, div3(add3(add3(add3(vtx(c, i - 1, j - 1, 0), vtx(c, i, j - 1, 0)), vtx(c, i - 1, j, 0)), vtx(c, i, j, 0)), 4.0));
                SETD(i, j, DK - 1, div3(add3(add3(add3(vtx(c, i - 1, j - 1, NK - 1), vtx(c, i, j - 1, NK - 1)),
                     vtx(c, i - 1, j, NK - 1)), vtx(c, i, j, NK - 1)), 4.0));
            }
    } else {
        /* 2-D and axisymmetric: src/mesh.cpp:373-435; 1-D: :438-475 */
        v3* nodes = (v3*)calloc((size_t)(DI * DJ), sizeof(v3));
#define ND(i, j) nodes[(long)(i) * DJ + (j)]
        if (c->p.topology == TBO_TOPO_1D) {
            for (long i = 1; i < DI - 1; i++) ND(i, 0) = div3(add3(vtx(c, i - 1, 0, 0), vtx(c, i, 0, 0)), 2.0);
            ND(0, 0) = vtx(c, 0, 0, 0);
            ND(DI - 1, 0) = vtx(c, NI - 1, 0, 0);
            for (long i = 0; i < DI; i++) {
                double xx = ND(i, 0).x;
                double* d;
                d = DN(i, 0, 0); d[0] = xx; d[1] = -0.5; d[2] = -0.5;
                d = DN(i, 1, 0); d[0] = xx; d[1] = 0.5; d[2] = -0.5;
                d = DN(i, 0, 1); d[0] = xx; d[1] = -0.5; d[2] = 0.5;
                d = DN(i, 1, 1); d[0] = xx; d[1] = 0.5; d[2] = 0.5;
            }
        } else {
            for (long i = 1; i < DI - 1; i++)
                for (long j = 1; j < DJ - 1; j++)
                    ND(i, j) = div3(add3(add3(add3(vtx(c, i - 1, j - 1, 0), vtx(c, i, j - 1, 0)), vtx(c, i, j, 0)), vtx(c, i - 1, j, 0)), 4.0);
            ND(0, 0) = vtx(c, 0, 0, 0);
            ND(0, DJ - 1) = vtx(c, 0, NJ - 1, 0);
            ND(DI - 1, 0) = vtx(c, NI - 1, 0, 0);
            ND(DI - 1, DJ - 1) = vtx(c, NI - 1, NJ - 1, 0);
            for (long i = 1; i < DI - 1; i++) {
                ND(i, 0) = div3(add3(vtx(c, i - 1, 0, 0), vtx(c, i, 0, 0)), 2.0);
                ND(i, DJ - 1) = div3(add3(vtx(c, i - 1, NJ - 1, 0), vtx(c, i, NJ - 1, 0)), 2.0);
            }
            for (long j = 1; j < DJ - 1; j++) {
                ND(0, j) = div3(add3(vtx(c, 0, j - 1, 0), vtx(c, 0, j, 0)), 2.0);
                ND(DI - 1, j) = div3(add3(vtx(c, NI - 1, j - 1, 0), vtx(c, NI - 1, j, 0)), 2.0);
            }
            if (c->p.topology == TBO_TOPO_AXI) {
                c->wedge_angle = 1.0 * M_PI / 180.0;
                double w = c->wedge_angle;
                for (long i = 0; i < DI; i++)
                    for (long j = 0; j < DJ; j++) {
                        double* d0 = DN(i, j, 0);
                        double* d1 = DN(i, j, 1);
                        d0[0] = ND(i, j).x; d1[0] = ND(i, j).x;
                        d0[1] = ND(i, j).y * cos(-w / 2.0); d1[1] = ND(i, j).y * cos(+w / 2.0);
                        d0[2] = ND(i, j).y * sin(-w / 2.0); d1[2] = ND(i, j).y * sin(+w / 2.0);
                    }
            } else {
                for (long i = 0; i < DI; i++)
                    for (long j = 0; j < DJ; j++) {
                        double* d0 = DN(i, j, 0);
                        double* d1 = DN(i, j, 1);
                        d0[0] = ND(i, j).x; d1[0] = ND(i, j).x;
                        d0[1] = ND(i, j).y; d1[1] = ND(i, j).y;
                        d0[2] = -1.0 / 2.0; d1[2] = +1.0 / 2.0;
                    }
            }
        }
#undef ND
        free(nodes);
    }
    /* faces: src/mesh.cpp:116-163 */
    c->sI = dalloc(DI * NJ * NK * 3); c->cI = dalloc(DI * NJ * NK * 3);
    c->sJ = dalloc(NI * DJ * NK * 3); c->cJ = dalloc(NI * DJ * NK * 3);
    c->sK = dalloc(NI * NJ * DK * 3); c->cK = dalloc(NI * NJ * DK * 3);
#define P3(ptr) mk3((ptr)[0], (ptr)[1], (ptr)[2])
    for (long i = 0; i < DI; i++)
        for (long j = 0; j < NJ; j++)
            for (long k = 0; k < NK; k++) {
                long f = ((i * NJ + j) * NK + k) * 3;
                surf_center(P3(DN(i, j, k)), P3(DN(i, j + 1, k)), P3(DN(i, j + 1, k + 1)), P3(DN(i, j, k + 1)), &c->sI[f], &c->cI[f]);
            }
    for (long i = 0; i < NI; i++)
        for (long j = 0; j < DJ; j++)
            for (long k = 0; k < NK; k++) {
                long f = ((i * DJ + j) * NK + k) * 3;
                surf_center(P3(DN(i, j, k)), P3(DN(i, j, k + 1)), P3(DN(i + 1, j, k + 1)), P3(DN(i + 1, j, k)), &c->sJ[f], &c->cJ[f]);
            }
    for (long i = 0; i < NI; i++)
        for (long j = 0; j < NJ; j++)
            for (long k = 0; k < DK; k++) {
                long f = ((i * NJ + j) * DK + k) * 3;
                surf_center(P3(DN(i, j, k)), P3(DN(i + 1, j, k)), P3(DN(i + 1, j + 1, k)), P3(DN(i, j + 1, k)), &c->sK[f], &c->cK[f]);
            }
    /* volumes: src/mesh.cpp:203-230, divergence theorem with x only */
    c->vol = dalloc(c->N);
    for (long i = 0; i < NI; i++)
        for (long j = 0; j < NJ; j++)
            for (long k = 0; k < NK; k++) {
                double vv = 0.0;
                vv += (-c->sI[FI(i, j, k)]) * c->cI[FI(i, j, k)];
                vv += (-c->sJ[FJ(i, j, k)]) * c->cJ[FJ(i, j, k)];
                vv += (-c->sK[FK(i, j, k)]) * c->cK[FK(i, j, k)];
                vv += c->sI[FI(i + 1, j, k)] * c->cI[FI(i + 1, j, k)];
                vv += c->sJ[FJ(i, j + 1, k)] * c->cJ[FJ(i, j + 1, k)];
                vv += c->sK[FK(i, j, k + 1)] * c->cK[FK(i, j, k + 1)];
                c->vol[IDX(c, i, j, k)] = vv;
            }
    /* radius / theta: include/mesh.hpp getRadius/getTheta */
    c->radius = dalloc(c->N);
    c->theta = dalloc(c->N);
    for (long n = 0; n < c->N; ++n) {
        c->radius[n] = sqrt(c->y[n] * c->y[n] + c->z[n] * c->z[n]);
        c->theta[n] = atan2(c->z[n], c->y[n]);
    }
    /* blockage gradient: src/mesh.cpp:571-581 -> math_utils.cpp:173-234, 395-462 */
    if (c->p.bfm_active && c->blockage) {
        c->gradb = dalloc(3 * c->N);
        for (long i = 0; i < NI; i++)
            for (long j = 0; j < NJ; j++)
                for (long k = 0; k < NK; k++) {
                    v3 surf[6] = {neg3(P3(&c->sI[FI(i, j, k)])), P3(&c->sI[FI(i + 1, j, k)]),
                                  P3(&c->sJ[FJ(i, j + 1, k)]), neg3(P3(&c->sJ[FJ(i, j, k)])),
                                  neg3(P3(&c->sK[FK(i, j, k)])), P3(&c->sK[FK(i, j, k + 1)])};
                    v3 cen[6] = {P3(&c->cI[FI(i, j, k)]), P3(&c->cI[FI(i + 1, j, k)]),
                                 P3(&c->cJ[FJ(i, j + 1, k)]), P3(&c->cJ[FJ(i, j, k)]),
                                 P3(&c->cK[FK(i, j, k)]), P3(&c->cK[FK(i, j, k + 1)])};
                    /* neighbour in each direction: W, E, N, S, B, T */
                    long nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j + 1, k}, {i, j - 1, k}, {i, j, k - 1}, {i, j, k + 1}};
                    int edge[6] = {i == 0, i == NI - 1, j == NJ - 1, j == 0, k == 0, k == NK - 1};
                    long id0 = IDX(c, i, j, k);
                    v3 p0 = vtx(c, i, j, k);
                    double f0 = c->blockage[id0];
                    v3 gsum = mk3(0.0, 0.0, 0.0);
                    for (int d = 0; d < 6; ++d) {
                        double val;
                        if (edge[d]) val = f0;
                        else {
                            long id1 = IDX(c, nb[d][0], nb[d][1], nb[d][2]);
                            v3 p1 = vtx(c, nb[d][0], nb[d][1], nb[d][2]);
                            double f1 = c->blockage[id1];
                            double d0 = mag3(sub3(cen[d], p0));
                            double d1 = mag3(sub3(cen[d], p1));
                            val = f0 + (f1 - f0) * d0 / (d0 + d1);
                        }
                        gsum = add3(gsum, mul3(surf[d], val));
                    }
                    v3 gr = div3(gsum, c->vol[id0]);
                    c->gradb[3 * id0 + 0] = gr.x; c->gradb[3 * id0 + 1] = gr.y; c->gradb[3 * id0 + 2] = gr.z;
                }
    }
#undef DN
#undef SETD
}

int tbo_geometry(const tbo_case* c, int which, double* out) {
    const long NI = c->p.ni, NJ = c->p.nj, NK = c->p.nk;
    long nI = (NI + 1) * NJ * NK * 3, nJ = NI * (NJ + 1) * NK * 3, nK = NI * NJ * (NK + 1) * 3;
    switch (which) {
        case 0: memcpy(out, c->sI, nI * 8); return 0;
        case 1: memcpy(out, c->sJ, nJ * 8); return 0;
        case 2: memcpy(out, c->sK, nK * 8); return 0;
        case 3: memcpy(out, c->cI, nI * 8); return 0;
        case 4: memcpy(out, c->cJ, nJ * 8); return 0;
        case 5: memcpy(out, c->cK, nK * 8); return 0;
        case 6: memcpy(out, c->vol, c->N * 8); return 0;
        case 7: memcpy(out, c->dual, (NI + 1) * (NJ + 1) * (NK + 1) * 3 * 8); return 0;
        case 8: if (!c->gradb) return -1; memcpy(out, c->gradb, 3 * c->N * 8); return 0;
        default: return -1;
    }
}

/* ------------------------------------------------------------------ BFM source */
typedef struct {
    double radius, cs, sn, omega, pitch, stag_pitch, flow_angle, dev, ntan, blockage;
    int stalled;
    v3 rel_cyl, rel_cart, inv_dir_cyl, visc_dir_cyl;
} flow_state;

/* src/source_bfm_base.cpp:152-166 */
static double deviation_angle(v3 rel, v3 nc) {
    v3 n = div3(nc, mag3(nc));
    v3 vm = mk3(rel.x, rel.y, 0.0);
    v3 sp = mk3(-vm.y, vm.x, 0.0);
    sp = div3(sp, mag3(sp));
    v3 vin = sub3(rel, mul3(sp, dot3(rel, sp)));
    double nv = dot3(vin, n);
    return -asin(nv / mag3(vin));
}
/* src/source_bfm_base.cpp:168-199; stalled: the camber normal itself once the deviation
 * exceeds the stall deviation (STALLED_BFM_ACTIVE) */
static v3 inviscid_dir(v3 rel, v3 nc, int stalled) {
    v3 rd = nrm3(rel), nd = nrm3(nc);
    v3 ver = stalled ? nd : nrm3(sub3(nd, mul3(rd, dot3(nd, rd))));
    v3 vm = mk3(rel.x, rel.y, 0.0);
    v3 sd = nrm3(mk3(vm.y, -vm.x, 0.0));
    return nrm3(sub3(ver, mul3(sd, dot3(ver, sd))));
}
/* Config::getRotationalSpeedScalingFactor(currentTime) (config.cpp:695-718): the constant
 * factor, or during [t0, t1] with ACCELERATION_ACTIVE the linear ramp s0 -> s1 */
static double omega_scale(const tbo_case* c, double t) {
    if (!c->p.accel_active) return c->p.omega_scaling;
    if (t < c->p.acc_t0 || t > c->p.acc_t1) return c->p.omega_scaling;
    return c->p.acc_s0 + (c->p.acc_s1 - c->p.acc_s0) * (t - c->p.acc_t0) / (c->p.acc_t1 - c->p.acc_t0);
}

/* SourceBFMBase::computePerturbationFactor (source_bfm_base.cpp:215-243): a Gaussian in
 * space about the centre (zero beyond three radii) and in time about the middle of the
 * perturbation window (sigma a third of its duration) */
static double perturbation_factor(const tbo_case* c, long idx, double t) {
    if (!c->p.pert_active) return 0.0;
    if (t < c->p.pert_start || t > c->p.pert_start + c->p.pert_duration) return 0.0;
    double vmax = c->p.pert_scale;
    v3 d = sub3(mk3(c->x[idx], c->y[idx], c->z[idx]), mk3(c->p.pert_center[0], c->p.pert_center[1], c->p.pert_center[2]));
    double distance = mag3(d);
    if (distance > 3.0 * c->p.pert_radius) return 0.0;
    double space = exp(-(distance * distance) / (c->p.pert_radius * c->p.pert_radius));
    double peak = c->p.pert_start + c->p.pert_duration / 2.0;
    double sigma = c->p.pert_duration / 3.0;
    double time = exp(-(pow(t - peak, 2)) / pow(sigma, 2));
    return vmax * space * time;
}

/* src/source_bfm_base.cpp:104-149 */
static flow_state compute_flow_state(const tbo_case* c, long idx, st5 W) {
    flow_state s;
    s.radius = c->radius[idx];
    double th = c->theta[idx];
    s.cs = cos(th); s.sn = sin(th);
    v3 vc = mk3(W.v[1], W.v[2], W.v[3]);
    v3 vcy = cyl_from_cart(vc, s.cs, s.sn);
    s.omega = c->rpm[idx] * 2 * M_PI / 60;
    s.omega *= omega_scale(c, c->t_phys);
    v3 drag = mk3(0, 0, s.omega * s.radius);
    s.rel_cyl = sub3(vcy, drag);
    s.rel_cart = cart_from_cyl(s.rel_cyl, s.cs, s.sn);
    double vmer = sqrt(vcy.x * vcy.x + vcy.y * vcy.y);
    double nb = c->nblades[idx];
    s.pitch = 2.0 * M_PI * s.radius / nb;
    v3 nc = mk3(c->nax[idx], c->nrad[idx], c->ntan[idx]);
    s.ntan = c->ntan[idx];
    double metal = c->metal ? c->metal[idx] : 0.0;
    s.stag_pitch = s.pitch * cos(metal);
    s.flow_angle = atan2(s.rel_cyl.z, vmer);
    s.dev = deviation_angle(s.rel_cyl, nc);
    s.stalled = c->p.stall_active && s.dev > (c->stall ? c->stall[idx] : 0.0);
    s.blockage = c->blockage ? c->blockage[idx] : 1.0;
    s.inv_dir_cyl = inviscid_dir(s.rel_cyl, nc, s.stalled);
    s.visc_dir_cyl = neg3(nrm3(s.rel_cyl));
    return s;
}

static st5 scale5(st5 a, double s) { for (int v = 0; v < 5; ++v) a.v[v] *= s; return a; }

/* SourceBFMThollet::computeCompressibilityCorrection (source_bfm_thollet.cpp:50-71):
 * relative Mach with the sound speed of computeSoundSpeed_rho_u_et, min(3, K') */
static double thollet_kmach(const tbo_case* c, v3 rel_cyl, st5 W) {
    double a = f_a(c, W.v[0], mk3(W.v[1], W.v[2], W.v[3]), W.v[4]);
    double m = mag3(rel_cyl) / a;
    if (fabs(m - 1.0) < 1E-03) m = 0.999;
    double kp = 0.0;
    if (m < 1.0) kp = 1.0 / sqrt(1.0 - m * m);
    else kp = 2.0 / M_PI / sqrt(m * m - 1.0);
    return (kp < 3.0) ? kp : 3.0;   /* std::min(3.0, kp) */
}

/* src/source_bfm_base.cpp:25-84 + source_bfm_hall.cpp:3-34 + source_bfm_lift_drag.cpp:3-88 */
static st5 bfm_total_source(tbo_case* c, long i, long j, long k, st5 W, int write_side) {
    long idx = IDX(c, i, j, k);
    st5 src = {{0, 0, 0, 0, 0}};
    st5 blk = {{0, 0, 0, 0, 0}};
    if (c->p.blockage_active) {
        v3 u = mk3(W.v[1], W.v[2], W.v[3]);
        double ht = f_ht(c, W.v[0], u, W.v[4]);
        double b = c->blockage[idx];
        v3 gb = mk3(c->gradb[3 * idx], c->gradb[3 * idx + 1], c->gradb[3 * idx + 2]);
        double V = c->vol[idx];
        double common = -1.0 / b * W.v[0] * dot3(u, gb);
        blk.v[0] = common;
        blk.v[1] = common * u.x;
        blk.v[2] = common * u.y;
        blk.v[3] = common * u.z;
        blk.v[4] = common * ht;
        blk = scale5(blk, V);
    }
    double nb = c->nblades ? c->nblades[idx] : 0.0;
    if (nb == 0) return blk;
    st5 body = {{0, 0, 0, 0, 0}};
    v3 finv = mk3(0, 0, 0), fvisc = mk3(0, 0, 0);
    double dev = 0.0;
    double V = c->vol[idx];
    if (c->p.bfm_model == TBO_BFM_HALL) {
        flow_state s = compute_flow_state(c, idx, W);
        double fm = dot3(s.rel_cyl, s.rel_cyl) * M_PI * s.dev / s.pitch / fabs(s.ntan);
        v3 fcyl = mul3(s.inv_dir_cyl, fm);
        v3 fcar = cart_from_cyl(fcyl, s.cs, s.sn);
        finv = fcar;
        st5 sv = {{0, fcar.x, fcar.y, fcar.z, fcyl.z * s.omega * s.radius}};
        body = scale5(scale5(sv, V), W.v[0]);
        dev = s.dev;
    } else if (c->p.bfm_model == TBO_BFM_LIFT_DRAG) {
        flow_state s = compute_flow_state(c, idx, W);
        double sl_te = c->slen[IDX(c, c->p.te_index, j, k)];
        double sl_le = c->slen[IDX(c, c->p.le_index, j, k)];
        double chord = sl_te - sl_le;
        double solidity = chord / s.pitch;
        double w = mag3(s.rel_cart);
        double beta0 = c->beta0[idx];
        double sign = (s.ntan >= 0.0) ? 1.0 : -1.0;
        double fm = 0.5 * w * w * 2.0 * M_PI / s.stag_pitch * sign * (-s.flow_angle + beta0);
        v3 fcyl = mul3(s.inv_dir_cyl, fm);
        v3 fcar = cart_from_cyl(fcyl, s.cs, s.sn);
        finv = fcar;
        st5 si = {{0, fcar.x, fcar.y, fcar.z, fcyl.z * s.omega * s.radius}};
        si = scale5(scale5(si, V), W.v[0]);
        double kp0 = c->kp[idx];
        double dd = s.flow_angle - beta0;
        double Kp = kp0 + 2.0 * M_PI * solidity * (dd * dd);
        double fm2 = Kp * w * w / s.stag_pitch;
        v3 vcyl = mul3(s.visc_dir_cyl, fm2);
        v3 vcar = cart_from_cyl(vcyl, s.cs, s.sn);
        fvisc = vcar;
        st5 sv = {{0, vcar.x, vcar.y, vcar.z, vcyl.z * s.omega * s.radius}};
        sv = scale5(scale5(sv, V), W.v[0]);
        for (int v = 0; v < 5; ++v) body.v[v] = si.v[v] + sv.v[v];
        dev = s.dev;
    }
    else if (c->p.bfm_model == TBO_BFM_THOLLET) {
        flow_state s = compute_flow_state(c, idx, W);
        /* inviscid component (source_bfm_thollet.cpp:23-48) */
        double kmach = thollet_kmach(c, s.rel_cyl, W);
        double fm = c->p.kn * kmach * dot3(s.rel_cyl, s.rel_cyl) * M_PI * s.dev / s.pitch / fabs(s.ntan) / s.blockage;
        v3 fcyl = mul3(s.inv_dir_cyl, fm);
        v3 fcar = cart_from_cyl(fcyl, s.cs, s.sn);
        finv = fcar;
        st5 si = {{0, fcar.x, fcar.y, fcar.z, fcyl.z * s.omega * s.radius}};
        si = scale5(scale5(si, V), W.v[0]);
        /* viscous component (:73-138): none once stalled, the side output keeping its
         * previous value */
        st5 sv = {{0, 0, 0, 0, 0}};
        int keep_visc = 0;
        if (s.stalled) {
            keep_visc = 1;
        } else {
            double stwl = c->slen[idx];
            if (stwl < 1E-06) {
                if (i + 1 < c->p.ni) stwl = (c->slen[IDX(c, i + 1, j, k)] + c->slen[idx]) / 2.0;
                if (stwl < 1E-06) stwl = 1E-06;
            }
            double rex = mag3(s.rel_cyl) * stwl / c->p.nu;
            double cf = 0.0592 * pow(rex, -0.2);
            double fm2;
            if (!c->p.thollet_off_design) {
                fm2 = dot3(s.rel_cyl, s.rel_cyl) / (s.pitch * s.blockage * fabs(s.ntan)) * (c->p.kf * cf);
            } else {
                double pivot = c->pivot[idx];
                double km = thollet_kmach(c, s.rel_cyl, W);
                fm2 = dot3(s.rel_cyl, s.rel_cyl) / (s.pitch * s.blockage * fabs(s.ntan)) *
                      (c->p.kf * cf + M_PI * km * pow(fabs(s.dev - pivot), 2.0 * c->p.kd));
            }
            v3 vcyl = mul3(s.visc_dir_cyl, fm2);
            v3 vcar = cart_from_cyl(vcyl, s.cs, s.sn);
            fvisc = vcar;
            st5 t = {{0, vcar.x, vcar.y, vcar.z, vcyl.z * s.omega * s.radius}};
            sv = scale5(scale5(t, V), W.v[0]);
        }
        for (int v = 0; v < 5; ++v) body.v[v] = si.v[v] + sv.v[v];
        dev = s.dev;
        if (keep_visc && write_side) fvisc = mk3(c->fvisc[3 * idx], c->fvisc[3 * idx + 1], c->fvisc[3 * idx + 2]);
    }
    if (write_side) {
        c->finv[3 * idx] = finv.x; c->finv[3 * idx + 1] = finv.y; c->finv[3 * idx + 2] = finv.z;
        c->fvisc[3 * idx] = fvisc.x; c->fvisc[3 * idx + 1] = fvisc.y; c->fvisc[3 * idx + 2] = fvisc.z;
        c->dev[idx] = dev;
    }
    double pf = perturbation_factor(c, idx, c->t_phys);
    for (int v = 0; v < 5; ++v) src.v[v] = blk.v[v] + body.v[v] * (1.0 + pf);
    return src;
}

void tbo_bfm_source(tbo_case* c, long i, long j, long k, const double* w, double* out,
                    double* finv, double* fvisc, double* dev) {
    st5 W;
    memcpy(W.v, w, 40);
    /* a single evaluation on fresh side-output arrays (a stalled node leaves its viscous
     * force untouched, i.e. zero here; in a run it keeps the previous iteration's) */
    long id0 = IDX(c, i, j, k);
    for (int q = 0; q < 3; ++q) c->finv[3 * id0 + q] = c->fvisc[3 * id0 + q] = 0.0;
    c->dev[id0] = 0.0;
    st5 s = bfm_total_source(c, i, j, k, W, 1);
    memcpy(out, s.v, 40);
    long idx = IDX(c, i, j, k);
    memcpy(finv, &c->finv[3 * idx], 24);
    memcpy(fvisc, &c->fvisc[3 * idx], 24);
    *dev = c->dev[idx];
}

/* solver.cpp:3074-3145 */
static st5 gong_source(const tbo_case* c, double omega, long i, long j, long k, double V, const double* U) {
    st5 z = {{0, 0, 0, 0, 0}};
    if (fabs(omega) < 1E-3) return z;
    const long NK = c->p.nk;
    st5 U0, U1, U2;
    if (omega > 0) {
        if (k == 0) { U0 = ld(c, U, IDX(c, i, j, k)); U1 = ld(c, U, IDX(c, i, j, NK - 2)); U2 = ld(c, U, IDX(c, i, j, NK - 3)); }
        else if (k == 1) { U0 = ld(c, U, IDX(c, i, j, k)); U1 = ld(c, U, IDX(c, i, j, 0)); U2 = ld(c, U, IDX(c, i, j, NK - 2)); }
        else { U0 = ld(c, U, IDX(c, i, j, k)); U1 = ld(c, U, IDX(c, i, j, k - 1)); U2 = ld(c, U, IDX(c, i, j, k - 2)); }
    } else {
        if (k == NK - 1) { U0 = ld(c, U, IDX(c, i, j, 0)); U1 = ld(c, U, IDX(c, i, j, 1)); U2 = ld(c, U, IDX(c, i, j, 2)); }
        else if (k == NK - 2) { U0 = ld(c, U, IDX(c, i, j, k)); U1 = ld(c, U, IDX(c, i, j, 0)); U2 = ld(c, U, IDX(c, i, j, 1)); }
        else { U0 = ld(c, U, IDX(c, i, j, k)); U1 = ld(c, U, IDX(c, i, j, k + 1)); U2 = ld(c, U, IDX(c, i, j, k + 2)); }
    }
    double dth = (2 * M_PI / (double)(NK - 1));
    st5 s;
    for (int v = 0; v < 5; ++v) {
        double d;
        if (omega > 0.0) d = (U0.v[v] * 3.0 - U1.v[v] * 4.0 + U2.v[v]) / (2 * dth);
        else d = (U0.v[v] * (-3.0) + U1.v[v] * 4.0 - U2.v[v]) / (2 * dth);
        s.v[v] = -d * omega;
    }
    return scale5(s, V);
}

/* solver.cpp:2790-2897; which: bit 1 geometric, 2 Gong, 4 body force (7 in the solver,
 * single terms for the tests that pin each) */
static void source_terms(tbo_case* c, const double* U, double* R, int which);
static void source_residuals(tbo_case* c, const double* U, double* R) { source_terms(c, U, R, 7); }
static void source_terms(tbo_case* c, const double* U, double* R, int which) {
    const int topo = c->p.topology;
    if (topo == TBO_TOPO_3D && !c->p.bfm_active) return;
    const long NI = c->p.ni, NJ = c->p.nj, NK = c->p.nk, N = c->N;
#pragma omp parallel for collapse(2) schedule(static)
    for (long i = 0; i < NI; i++)
        for (long j = 0; j < NJ; j++)
            for (long k = 0; k < NK; k++) {
                long idx = IDX(c, i, j, k);
                st5 W = prim(ld(c, U, idx));
                double V = c->vol[idx];
                v3 u = mk3(W.v[1], W.v[2], W.v[3]);
                double p = f_p(c, W.v[0], u, W.v[4]);
                double r = c->radius[idx];
                double th = c->theta[idx];
                int geo = 0, gong = 0;
                if (topo == TBO_TOPO_3D && c->p.bfm_active) geo = gong = (c->present[idx] > 0.0);
                else if (topo == TBO_TOPO_AXI) geo = 1;
                geo = geo && (which & 1);
                gong = gong && (which & 2);
                if (geo) {
                    double cs = cos(th), sn = sin(th);
                    v3 vcy = cyl_from_cart(u, cs, sn);
                    v3 scy;
                    st5 sg;
                    if (topo == TBO_TOPO_AXI) {
                        long fk = FK(i, j, 0);
                        double areaK = mag3(mk3(c->sK[fk], c->sK[fk + 1], c->sK[fk + 2]));
                        double gf = areaK * c->wedge_angle;
                        scy = mk3(0.0, (+W.v[0] * vcy.z * vcy.z + p) * gf, -W.v[0] * vcy.y * vcy.z * gf);
                        v3 sca = cart_from_cyl(scy, cs, sn);
                        st5 t = {{0.0, sca.x, sca.y, sca.z, 0.0}};
                        sg = t;
                    } else {
                        scy = mk3(0.0, (+W.v[0] * vcy.z * vcy.z + p) / r, -W.v[0] * vcy.y * vcy.z / r);
                        v3 sca = cart_from_cyl(scy, cs, sn);
                        st5 t = {{0.0, sca.x, sca.y, sca.z, 0.0}};
                        sg = scale5(t, V);
                    }
                    for (int v = 0; v < 5; ++v) R[v * N + idx] -= sg.v[v];
                }
                if (gong) {
                    double om = c->rpm[idx] * 2 * M_PI / 60;
                    om *= omega_scale(c, c->t_phys);
                    st5 gs = gong_source(c, om, i, j, k, V, U);
                    for (int v = 0; v < 5; ++v) R[v * N + idx] -= gs.v[v];
                }
                if (c->p.bfm_active && (which & 4)) {
                    v3 gb = mk3(c->gradb[3 * idx], c->gradb[3 * idx + 1], c->gradb[3 * idx + 2]);
                    if (mag3(gb) > 1E-10) {
                        st5 bs = bfm_total_source(c, i, j, k, W, 1);
                        for (int v = 0; v < 5; ++v) R[v * N + idx] -= bs.v[v];
                    }
                }
            }
}

/* ------------------------------------------------------------ residual assembly */
static inline void radd(double* R, long N, long idx, st5 F, double s) {
    for (int v = 0; v < 5; ++v) R[v * N + idx] += F.v[v] * s;
}
static inline void rsub(double* R, long N, long idx, st5 F, double s) {
    for (int v = 0; v < 5; ++v) R[v * N + idx] -= F.v[v] * s;
}

/* solver.cpp:1865-2040, for one line of faces along direction d */
static void face_line(tbo_case* c, int d, long a, long b, const double* U, double* R, long it) {
    const long NI = c->p.ni, NJ = c->p.nj, NK = c->p.nk, N = c->N;
    const long nf = (d == 0) ? NI + 1 : (d == 1) ? NJ + 1 : NK + 1;
    const long stop = nf - 1;
    const double* Sarr = (d == 0) ? c->sI : (d == 1) ? c->sJ : c->sK;
    const int* map = (d == 0) ? c->mapI : (d == 1) ? c->mapJ : c->mapK;
    const double ang = c->per_angle_rad;
    for (long f = 0; f < nf; ++f) {
        long fi, fj, fk;
        if (d == 0) { fi = f; fj = a; fk = b; }
        else if (d == 1) { fi = a; fj = f; fk = b; }
        else { fi = a; fj = b; fk = f; }
        if (d == 2 && c->p.bfm_active) {
            if (c->present[IDX(c, fi, fj, 0)] > 0.0) continue;   /* :1903-1909 */
        }
        long fidx;
        if (d == 0) fidx = ((fi * NJ + fj) * NK + fk) * 3;
        else if (d == 1) fidx = ((fi * (NJ + 1) + fj) * NK + fk) * 3;
        else fidx = ((fi * NJ + fj) * (NK + 1) + fk) * 3;
        v3 S = mk3(Sarr[fidx], Sarr[fidx + 1], Sarr[fidx + 2]);
        /* map index: 2 layers along d; the other two indices are the face's */
        long m0 = (d == 0) ? 0 : fi, m1 = (d == 1) ? 0 : fj, m2 = (d == 2) ? 0 : fk;
        long mdim1 = (d == 1) ? 2 : NJ, mdim2 = (d == 2) ? 2 : NK;
        if (f == 0) {
            st5 Ui = ld(c, U, IDX(c, fi, fj, fk));
            v3 Sb = neg3(S);
            int pidx = map[(m0 * mdim1 + m1) * mdim2 + m2];
            st5 F = bc_flux(c, pidx, Ui, Sb, fi, fj, fk, it);
            radd(R, N, IDX(c, fi, fj, fk), F, mag3(Sb));
        } else if (f == stop) {
            long li = fi - (d == 0), lj = fj - (d == 1), lk = fk - (d == 2);
            st5 Ui = ld(c, U, IDX(c, li, lj, lk));
            if (d == 0) m0 = 1; else if (d == 1) m1 = 1; else m2 = 1;
            int pidx = map[(m0 * mdim1 + m1) * mdim2 + m2];
            st5 F = bc_flux(c, pidx, Ui, S, fi, fj, fk, it);
            radd(R, N, IDX(c, li, lj, lk), F, mag3(S));
        } else {
            long di = (d == 0), dj = (d == 1), dk = (d == 2);
            st5 Ul = ld(c, U, IDX(c, fi - di, fj - dj, fk - dk));
            st5 Ur = ld(c, U, IDX(c, fi, fj, fk));
            st5 Ull, Urr;
            if (d == 2 && c->periodic) {
                if (f == 1) {
                    Ull = rotx(ld(c, U, IDX(c, fi, fj, nf - 3)), cos(-ang), sin(-ang));
                    Urr = ld(c, U, IDX(c, fi, fj, fk + 1));
                } else if (f == stop - 1) {
                    Ull = ld(c, U, IDX(c, fi, fj, fk - 2));
                    Urr = rotx(ld(c, U, IDX(c, fi, fj, 1)), cos(ang), sin(ang));
                } else {
                    Ull = ld(c, U, IDX(c, fi - 2 * di, fj - 2 * dj, fk - 2 * dk));
                    Urr = ld(c, U, IDX(c, fi + di, fj + dj, fk + dk));
                }
            } else {
                if (f == 1) {
                    for (int v = 0; v < 5; ++v) Ull.v[v] = Ul.v[v] - (Ur.v[v] - Ul.v[v]);
                    Urr = ld(c, U, IDX(c, fi + di, fj + dj, fk + dk));
                } else if (f == stop - 1) {
                    Ull = ld(c, U, IDX(c, fi - 2 * di, fj - 2 * dj, fk - 2 * dk));
                    for (int v = 0; v < 5; ++v) Urr.v[v] = Ur.v[v] + (Ur.v[v] - Ul.v[v]);
                } else {
                    Ull = ld(c, U, IDX(c, fi - 2 * di, fj - 2 * dj, fk - 2 * dk));
                    Urr = ld(c, U, IDX(c, fi + di, fj + dj, fk + dk));
                }
            }
            st5 F = face_flux(c, Ull, Ul, Ur, Urr, S);
            double sm = mag3(S);
            radd(R, N, IDX(c, fi - di, fj - dj, fk - dk), F, sm);
            rsub(R, N, IDX(c, fi, fj, fk), F, sm);
        }
    }
}

static void advection_residuals(tbo_case* c, int d, const double* U, double* R, long it) {
    const long NI = c->p.ni, NJ = c->p.nj, NK = c->p.nk;
    long A = (d == 0) ? NJ : NI;
    long B = (d == 2) ? NJ : NK;
#pragma omp parallel for collapse(2) schedule(static)
    for (long a = 0; a < A; ++a)
        for (long b = 0; b < B; ++b) face_line(c, d, a, b, U, R, it);
}

/* solver.cpp:1836-1853 */
static void periodic_residuals(tbo_case* c, double* R) {
    const long NI = c->p.ni, NJ = c->p.nj, NK = c->p.nk, N = c->N;
    double a = c->per_angle_rad;
    for (long i = 0; i < NI; i++)
        for (long j = 0; j < NJ; j++) {
            long i0 = IDX(c, i, j, 0), i1 = IDX(c, i, j, NK - 1);
            st5 R1 = ld(c, R, i0), R2 = ld(c, R, i1);
            st5 R2f = rotx(R2, cos(-a), sin(-a));
            st5 avg;
            for (int v = 0; v < 5; ++v) avg.v[v] = (R1.v[v] + R2f.v[v]) * 0.5;
            st5 back = rotx(avg, cos(+a), sin(+a));
            for (int v = 0; v < 5; ++v) { R[v * N + i0] = avg.v[v]; R[v * N + i1] = back.v[v]; }
        }
}
/* solver.cpp:2308-2327 */
static void periodic_solution(tbo_case* c, double* U) {
    const long NI = c->p.ni, NJ = c->p.nj, NK = c->p.nk, N = c->N;
    double a = c->per_angle_rad;
    for (long i = 0; i < NI; ++i)
        for (long j = 0; j < NJ; ++j) {
            long i0 = IDX(c, i, j, 0), i1 = IDX(c, i, j, NK - 1);
            st5 U1 = ld(c, U, i0), U2 = ld(c, U, i1);
            st5 U2r = rotx(U2, cos(-a), sin(-a));
            st5 avg;
            for (int v = 0; v < 5; ++v) avg.v[v] = (U1.v[v] + U2r.v[v]) * 0.5;
            st5 back = rotx(avg, cos(a), sin(a));
            for (int v = 0; v < 5; ++v) { U[v * N + i0] = avg.v[v]; U[v * N + i1] = back.v[v]; }
        }
}

/* Solver::updateRadialProfiles (solver.cpp:2688-2787): the hub pressure of each outlet
 * profile (radial equilibrium: kept; throttle: p0 of the first inlet + k_t m^2 with the
 * turbomachinery mass flow; mass flow: relaxed towards the target mass flow of the
 * patch's boundary), then the radial-equilibrium integration or a uniform profile. */
void tbo_turbo(tbo_case* c, const double* U, double* out);

/* Throttle coefficient of the Greitzer model: the last throttle patch's (solver.cpp:733-741) */
static double greitzer_kt(const tbo_case* c) {
    double kt = 0.0;
    for (int b = 0; b < c->npatches; ++b)
        if (c->patches[b].type == TBO_BC_THROTTLE) kt = c->patches[b].values[0];
    return kt;
}
/* FluidBase::computeSoundSpeed_p_T: ideal sqrt(gamma R T) (fluid_ideal.cpp), table a(p, T)
 * (fluid_real.cpp:422-424) */
static double f_a_p_T(const tbo_case* c, double p, double T) {
    if (c->p.fluid_real) return lut_interp(c, TBO_LUT_A_PT, p, T);
    return sqrt(c->p.gamma * c->p.r_gas * T);
}
/* GreitzerModel::computePlenumPressure (greitzer_model.cpp:36-56) */
static double greitzer_plenum_pressure(tbo_case* c, double mflow) {
    double min = mflow;
    double a = f_a_p_T(c, c->gz_p, c->gz_T);
    double np = c->gz_p + c->p.fixed_dt * a * a / c->p.greitzer_volume * (min - c->gz_mout);
    c->gz_p = np;
    double dp = c->gz_p - c->p.greitzer_ambient;
    c->gz_mout = (dp > 0.0) ? sqrt(dp / greitzer_kt(c)) : 0.0;
    return c->gz_p;
}

static void update_radial_profiles(tbo_case* c, const double* U) {
    double t[10];
    int need_mf = 0;
    for (int q = 0; q < c->nprof; ++q) need_mf |= (c->patches[c->prof_patch[q]].type != TBO_BC_RADIAL_EQ);
    need_mf |= c->p.greitzer;
    if (need_mf) tbo_turbo(c, U, t);   /* updateMassFlows + updateTurboPerformance */
    double gz_hub = 0.0;
    if (c->p.greitzer) gz_hub = greitzer_plenum_pressure(c, t[6]);
    for (int q = 0; q < c->nprof; ++q) {
        const tbo_patch* pa = &c->patches[c->prof_patch[q]];
        long n = c->prof_n[q];
        double* rho = dalloc(n);
        double* ut = dalloc(n);
        long i = (pa->i0 > 0) ? (pa->i0 - 1) : 0;
        long k = 0;
        for (long j = pa->j0; j < pa->j1 && k < n; j++) {
            long idx = IDX(c, i, j, 0);
            st5 W = prim(ld(c, U, idx));
            double th = c->theta[idx];
            v3 vcy = cyl_from_cart(mk3(W.v[1], W.v[2], W.v[3]), cos(th), sin(th));
            rho[k] = W.v[0];
            ut[k] = fabs(vcy.z);
            k++;
        }
        double hub = c->prof_p[q][0];
        if (c->p.greitzer) {
            hub = gz_hub;   /* every profile follows the plenum (solver.cpp:2695-2698, 2719-2721) */
        } else if (pa->type == TBO_BC_THROTTLE) {
            double kt = pa->values[0], mflow = t[6], pt_in = 0.0;
            for (int b = 0; b < c->npatches; ++b)
                if (c->patches[b].type == TBO_BC_INLET) { pt_in = c->patches[b].values[0]; break; }
            hub = pt_in + kt * mflow * mflow;
        } else if (pa->type == TBO_BC_MASS_FLOW) {
            double target = pa->values[0];
            /* BoundaryIndex of the patch: I_START (i = 0) or I_END (i = NI) */
            double cur = fabs(t[(pa->i0 == 0) ? 0 : 1]);
            if (c->p.topology == TBO_TOPO_AXI) cur *= 2.0 * M_PI / c->wedge_angle;
            else if (c->per_angle_deg != 0.0) cur *= 360.0 / c->per_angle_deg;
            double relax = c->p.mass_flow_relaxation;
            if (pa->nvalues >= 3 && pa->values[2] > 0.0) relax = pa->values[2];
            double maxc = c->p.mass_flow_max_change;
            double err = (cur - target) / target;
            double d = relax * err;
            d = (d < -maxc) ? -maxc : ((maxc < d) ? maxc : d);   /* std::clamp */
            double cp = c->prof_p[q][0];
            if (cp <= 0.0) cp = c->p.init_pressure > 0.0 ? c->p.init_pressure : 101325.0;
            hub = cp * (1.0 + d);
            hub = (hub < 100.0) ? 100.0 : hub;   /* std::max(hub, 100) */
        }
        if (c->prof_radial[q] && n > 1) tbo_radial_equilibrium(rho, ut, c->prof_r[q], (int)n, hub, c->prof_p[q]);
        else for (long m = 0; m < n; ++m) c->prof_p[q][m] = hub;
        free(rho);
        free(ut);
    }
}

/* ------------------------------------------------------------- viscous terms */
/* FluidIdeal::computeMolecularDynamicViscosity (fluid_ideal.cpp:224-231) */
static double f_mu(const tbo_case* c, double T) {
    if (c->p.mu_constant_model) return c->p.mu_constant;
    const double Tr = c->p.mu_t_ref, Sc = c->p.mu_s;
    return c->p.mu_ref * pow(T / Tr, 1.5) * ((Tr + Sc) / (T + Sc));
}

/* Solver::setMomentumSolutionOnViscousWalls with the wall velocity (solver.cpp:3147-3185),
 * as preprocessSolution applies it (solver.cpp:3189-3204) */
static void wall_momentum(const tbo_case* c, double* U) {
    const long N = c->N;
    for (long n = 0; n < N; ++n) {
        if (!c->nsw[n]) continue;
        const double rho = U[n];
        for (int e = 1; e <= 3; ++e) U[e * N + n] = rho * c->wallv[3 * n + e - 1];
    }
}

/* Solver::computeSolutionGradient (solver.cpp:3033-3058): velocity and temperature of
 * every node, then computeGradientsGreenGauss4 (math_utils.cpp:237-392): face values
 * interpolated by the distances of the two nodes to the face centre, the node value on
 * boundary faces, summed over the faces W, E, N, S, B, T and divided by the volume. */
static void solution_gradient(tbo_case* c, const double* U) {
    const long N = c->N, NI = c->p.ni, NJ = c->p.nj, NK = c->p.nk;
    double* f = (double*)malloc(sizeof(double) * 4 * N);
    const double k = (c->p.gamma - 1.0) / c->p.r_gas;
    for (long n = 0; n < N; ++n) {
        double rho = U[n];
        double ux = U[N + n] / rho, uy = U[2 * N + n] / rho, uz = U[3 * N + n] / rho, et = U[4 * N + n] / rho;
        f[n] = ux; f[N + n] = uy; f[2 * N + n] = uz;
        f[3 * N + n] = (et - (ux * ux + uy * uy + uz * uz) * 0.5) * k;
    }
#pragma omp parallel for collapse(2) schedule(static)
    for (long i = 0; i < NI; ++i)
        for (long j = 0; j < NJ; ++j)
            for (long kk = 0; kk < NK; ++kk) {
                const long n = IDX(c, i, j, kk);
                const v3 p0 = vtx(c, i, j, kk);
                /* neighbour, face centre and outward face vector of W E N S B T */
                long nb[6];
                const double* mid[6];
                v3 sf[6];
                nb[0] = (i == 0) ? -1 : IDX(c, i - 1, j, kk);        mid[0] = &c->cI[FI(i, j, kk)];
                nb[1] = (i == NI - 1) ? -1 : IDX(c, i + 1, j, kk);   mid[1] = &c->cI[FI(i + 1, j, kk)];
                nb[2] = (j == NJ - 1) ? -1 : IDX(c, i, j + 1, kk);   mid[2] = &c->cJ[FJ(i, j + 1, kk)];
                nb[3] = (j == 0) ? -1 : IDX(c, i, j - 1, kk);        mid[3] = &c->cJ[FJ(i, j, kk)];
                nb[4] = (kk == 0) ? -1 : IDX(c, i, j, kk - 1);       mid[4] = &c->cK[FK(i, j, kk)];
                nb[5] = (kk == NK - 1) ? -1 : IDX(c, i, j, kk + 1);  mid[5] = &c->cK[FK(i, j, kk + 1)];
                sf[0] = neg3(P3(&c->sI[FI(i, j, kk)]));
                sf[1] = P3(&c->sI[FI(i + 1, j, kk)]);
                sf[2] = P3(&c->sJ[FJ(i, j + 1, kk)]);
                sf[3] = neg3(P3(&c->sJ[FJ(i, j, kk)]));
                sf[4] = neg3(P3(&c->sK[FK(i, j, kk)]));
                sf[5] = P3(&c->sK[FK(i, j, kk + 1)]);
                v3 g[4] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                for (int q = 0; q < 6; ++q) {
                    double w = 0.0;
                    if (nb[q] >= 0) {
                        const v3 p1 = mk3(c->x[nb[q]], c->y[nb[q]], c->z[nb[q]]);
                        const v3 m = P3(mid[q]);
                        const double d0 = mag3(sub3(m, p0)), d1 = mag3(sub3(m, p1));
                        w = d0 / (d0 + d1);
                    }
                    for (int v = 0; v < 4; ++v) {
                        const double fc = f[v * N + n];
                        const double sv = (nb[q] < 0) ? fc : fc + (f[v * N + nb[q]] - fc) * w;
                        g[v] = add3(g[v], mul3(sf[q], sv));
                    }
                }
                const double iv = 1.0 / c->vol[n];
                for (int v = 0; v < 4; ++v) {
                    c->grad[(3 * v) * N + n] = g[v].x * iv;
                    c->grad[(3 * v + 1) * N + n] = g[v].y * iv;
                    c->grad[(3 * v + 2) * N + n] = g[v].z * iv;
                }
            }
    free(f);
}

void tbo_gradients(const tbo_case* c, double* out) {
    if (c->grad) memcpy(out, c->grad, sizeof(double) * 12 * c->N);
}

/* Solver::computeViscousFlux (solver.cpp:2173-2226) with no eddy viscosity
 * (turbulence_model_none.hpp:49-51) and computeViscousStressTensor (math_utils.cpp:593-608):
 * flux per unit area through the face direction S/|S| */
static st5 viscous_flux(const tbo_case* c, st5 U, v3 gx, v3 gy, v3 gz, v3 gt, v3 S) {
    st5 W = prim(U);
    v3 vel = mk3(W.v[1], W.v[2], W.v[3]);
    double T = f_T(c, W.v[0], vel, W.v[4]);
    double mu = f_mu(c, T) + 0.0;
    double kap = c->p.cp_visc * f_mu(c, T) / c->p.prandtl + 0.0;
    double lam = -2.0 / 3.0 * mu;
    double div = gx.x + gy.y + gz.z;
    double txx = lam * div + 2.0 * mu * gx.x;
    double tyy = lam * div + 2.0 * mu * gy.y;
    double tzz = lam * div + 2.0 * mu * gz.z;
    double txy = mu * (gx.y + gy.x), txz = mu * (gx.z + gz.x), tyz = mu * (gy.z + gz.y);
    v3 tX = mk3(txx, txy, txz), tY = mk3(txy, tyy, tyz), tZ = mk3(txz, tyz, tzz);
    double thx = dot3(tX, vel) + kap * gt.x;
    double thy = dot3(tY, vel) + kap * gt.y;
    double thz = dot3(tZ, vel) + kap * gt.z;
    v3 n = div3(S, mag3(S));
    st5 F = {{0.0, dot3(tX, n), dot3(tY, n), dot3(tZ, n), thx * n.x + thy * n.y + thz * n.z}};
    return F;
}

void tbo_viscous_flux(const tbo_case* c, const double* U, const double* g, const double* S, double* out) {
    st5 u;
    memcpy(u.v, U, 40);
    st5 F = viscous_flux(c, u, mk3(g[0], g[1], g[2]), mk3(g[3], g[4], g[5]), mk3(g[6], g[7], g[8]),
                         mk3(g[9], g[10], g[11]), mk3(S[0], S[1], S[2]));
    memcpy(out, F.v, 40);
}

/* Solver::computeViscousFluxResiduals (solver.cpp:2060-2170): internal faces only, state
 * and gradients averaged over the two nodes, flux*|S|*(-1) added to the left node and
 * subtracted from the right one; faces of a grid line in increasing order */
static void viscous_residuals(tbo_case* c, int d, const double* U, double* R) {
    const long NI = c->p.ni, NJ = c->p.nj, NK = c->p.nk, N = c->N;
    const long A = (d == 0) ? NJ : NI, B = (d == 2) ? NJ : NK, L = (d == 0) ? NI : ((d == 1) ? NJ : NK);
#pragma omp parallel for collapse(2) schedule(static)
    for (long a = 0; a < A; ++a)
        for (long b = 0; b < B; ++b)
            for (long f = 1; f < L; ++f) {
                long il, jl, kl, ir, jr, kr;
                v3 S;
                if (d == 0) { ir = f; jr = a; kr = b; il = f - 1; jl = a; kl = b; S = P3(&c->sI[FI(ir, jr, kr)]); }
                else if (d == 1) { ir = a; jr = f; kr = b; il = a; jl = f - 1; kl = b; S = P3(&c->sJ[FJ(ir, jr, kr)]); }
                else { ir = a; jr = b; kr = f; il = a; jl = b; kl = f - 1; S = P3(&c->sK[FK(ir, jr, kr)]); }
                const long l = IDX(c, il, jl, kl), r = IDX(c, ir, jr, kr);
                st5 Ua;
                for (int v = 0; v < 5; ++v) Ua.v[v] = (U[v * N + l] + U[v * N + r]) * 0.5;
                v3 g[4];
                for (int v = 0; v < 4; ++v) {
                    const double* G = c->grad + (long)(3 * v) * N;
                    g[v] = mul3(add3(mk3(G[l], G[N + l], G[2 * N + l]), mk3(G[r], G[N + r], G[2 * N + r])), 0.5);
                }
                st5 F = viscous_flux(c, Ua, g[0], g[1], g[2], g[3], S);
                const double m = mag3(S);
                for (int v = 0; v < 5; ++v) {
                    const double fm = F.v[v] * m * (-1.0);
                    R[v * N + l] += fm;
                    R[v * N + r] -= fm;
                }
            }
}

/* Solver::enforceNoSlipWallsOnResiduals (solver.cpp:1855-1862): momentum residual of a
 * wall node = its density residual times the zero wall effect */
static void noslip_residuals(const tbo_case* c, double* R) {
    const long N = c->N;
    for (long n = 0; n < N; ++n)
        if (c->nsw[n])
            for (int e = 1; e <= 3; ++e) R[e * N + n] = R[n] * 0.0;
}

/* Solver::computeResiduals (solver.cpp:1782-1833) */
static void compute_residuals(tbo_case* c, const double* U, double* R, long it) {
    memset(R, 0, sizeof(double) * 5 * c->N);
    advection_residuals(c, 0, U, R, it);
    if (c->p.topology != TBO_TOPO_1D) advection_residuals(c, 1, U, R, it);
    if (c->p.topology == TBO_TOPO_3D) advection_residuals(c, 2, U, R, it);
    if (c->p.viscous) {
        solution_gradient(c, U);
        viscous_residuals(c, 0, U, R);
        viscous_residuals(c, 1, U, R);
        if (c->p.topology == TBO_TOPO_3D) viscous_residuals(c, 2, U, R);
    }
    source_residuals(c, U, R);
    if (c->periodic) periodic_residuals(c, R);
    if (c->p.viscous) noslip_residuals(c, R);
}

/* solver.cpp:1531-1630 */
void tbo_timestep(tbo_case* c, const double* U, double* dt) {
    const long NI = c->p.ni, NJ = c->p.nj, NK = c->p.nk;
    const long DJ = NJ + 1, DK = NK + 1;
    if (c->p.dt_method == TBO_DT_FIXED) {
        for (long n = 0; n < c->N; ++n) dt[n] = c->p.fixed_dt;
        return;
    }
#pragma omp parallel for collapse(2) schedule(static)
    for (long i = 0; i < NI; i++)
        for (long j = 0; j < NJ; j++)
            for (long k = 0; k < NK; k++) {
                const double* p0 = &c->dual[((i * DJ + j) * DK + k) * 3];
                const double* pI = &c->dual[(((i + 1) * DJ + j) * DK + k) * 3];
                const double* pJ = &c->dual[((i * DJ + j + 1) * DK + k) * 3];
                const double* pK = &c->dual[((i * DJ + j) * DK + k + 1) * 3];
                v3 eI = sub3(P3(pI), P3(p0)), eJ = sub3(P3(pJ), P3(p0)), eK = sub3(P3(pK), P3(p0));
                double dsI = mag3(eI), dsJ = mag3(eJ), dsK = mag3(eK);
                v3 iD = div3(eI, dsI), jD = div3(eJ, dsJ), kD = div3(eK, dsK);
                long idx = IDX(c, i, j, k);
                st5 W = prim(ld(c, U, idx));
                v3 u = mk3(W.v[1], W.v[2], W.v[3]);
                double a = f_a(c, W.v[0], u, W.v[4]);
                double ui = fabs(dot3(u, iD)), uj = fabs(dot3(u, jD)), uk = fabs(dot3(u, kD));
                double lam = (ui + a) / dsI;
                if (c->p.topology != TBO_TOPO_1D) lam += (uj + a) / dsJ;
                if (c->p.topology == TBO_TOPO_3D) lam += (uk + a) / dsK;
                double mu = c->p.viscous ? f_mu(c, f_T(c, W.v[0], u, W.v[4])) : 0.0;
                double nu = mu / W.v[0], nut = 0.0 / W.v[0];
                double invH2 = 1.0 / (dsI * dsI);
                if (c->p.topology != TBO_TOPO_1D) invH2 += 1.0 / (dsJ * dsJ);
                if (c->p.topology == TBO_TOPO_3D) invH2 += 1.0 / (dsK * dsK);
                double lv = 4.0 * (nu + nut) * invH2;
                dt[idx] = c->p.cfl / (lam + lv);
            }
    if (c->p.dt_method == TBO_DT_GLOBAL) {
        double m = dt[0];
        for (long n = 1; n < c->N; ++n) if (dt[n] < m) m = dt[n];
        for (long n = 0; n < c->N; ++n) dt[n] = m;
    }
}

/* solver.cpp:2230-2249 */
static void update_solution(tbo_case* c, const double* U0, double* Un, const double* R, double alpha) {
    const long N = c->N;
#pragma omp parallel for schedule(static)
    for (long idx = 0; idx < N; ++idx) {
        double factor = alpha * c->dt[idx] / c->vol[idx];
        for (int v = 0; v < 5; ++v) Un[v * N + idx] = U0[v * N + idx] - R[v * N + idx] * factor;
    }
}

/* solver.cpp:1488-1508 */
static int log_norms(const tbo_case* c, const double* R, double* out) {
    for (int v = 0; v < 5; ++v) {
        double s = 0.0;
        for (long n = 0; n < c->N; ++n) s += R[v * c->N + n] * R[v * c->N + n];
        double nrm = sqrt(s);
        if (isnan(nrm) || isinf(nrm)) return -1 - v;
        if (nrm >= 2.2250738585072014e-308) out[v] = log10(nrm / (double)(c->p.ni * c->p.nj * c->p.nk));
        else out[v] = -16.0;
        if (out[v] < -16.0) out[v] = -16.0;
    }
    return 0;
}

/* solver.cpp:1633-1781 (ideal gas) */
void tbo_turbo(tbo_case* c, const double* U, double* out) {
    const long NI = c->p.ni, NJ = c->p.nj, NK = c->p.nk;
    double mf[6];
    /* boundary slices: I_START, I_END, J_START, J_END, K_START, K_END */
    for (int b = 0; b < 6; ++b) {
        double sum = 0.0;
        long n1, n2;
        if (b < 2) { n1 = NJ; n2 = NK; } else if (b < 4) { n1 = NI; n2 = NK; } else { n1 = NI; n2 = NJ; }
        for (long p = 0; p < n1; ++p)
            for (long q = 0; q < n2; ++q) {
                long f, node;
                if (b == 0) { f = FI(0, p, q); node = IDX(c, 0, p, q); }
                else if (b == 1) { f = FI(NI, p, q); node = IDX(c, NI - 1, p, q); }
                else if (b == 2) { f = (((long)p * (NJ + 1) + 0) * NK + q) * 3; node = IDX(c, p, 0, q); }
                else if (b == 3) { f = (((long)p * (NJ + 1) + NJ) * NK + q) * 3; node = IDX(c, p, NJ - 1, q); }
                else if (b == 4) { f = FK(p, q, 0); node = IDX(c, p, q, 0); }
                else { f = FK(p, q, NK); node = IDX(c, p, q, NK - 1); }
                const double* S = (b < 2) ? &c->sI[f] : (b < 4) ? &c->sJ[f] : &c->sK[f];
                sum += S[0] * U[1 * c->N + node] + S[1] * U[2 * c->N + node] + S[2] * U[3 * c->N + node];
            }
        mf[b] = sum;
        out[b] = sum;
    }
    int bi[2] = {0, 1};
    if (c->p.turbo_dir == 1) { bi[0] = 1; bi[1] = 0; }
    else if (c->p.turbo_dir == 2) { bi[0] = 2; bi[1] = 3; }
    else if (c->p.turbo_dir == 3) { bi[0] = 3; bi[1] = 2; }
    double m = 0.5 * (fabs(mf[bi[0]]) + fabs(mf[bi[1]]));
    if (c->p.topology == TBO_TOPO_AXI) m *= 2.0 * M_PI / c->wedge_angle;
    else if (c->per_angle_deg != 0.0) m *= 360.0 / c->per_angle_deg;
    out[6] = m;
    double ptm[2], ttm[2], htm[2], sm[2];
    for (int s = 0; s < 2; ++s) {
        int b = bi[s];
        double sp = 0.0, stt = 0.0, sht = 0.0, ssv = 0.0;
        long n1 = (b < 2) ? NJ : NI, n2 = NK;
        for (long p = 0; p < n1; ++p)
            for (long q = 0; q < n2; ++q) {
                long f, node;
                if (b == 0) { f = FI(0, p, q); node = IDX(c, 0, p, q); }
                else if (b == 1) { f = FI(NI, p, q); node = IDX(c, NI - 1, p, q); }
                else if (b == 2) { f = (((long)p * (NJ + 1) + 0) * NK + q) * 3; node = IDX(c, p, 0, q); }
                else { f = (((long)p * (NJ + 1) + NJ) * NK + q) * 3; node = IDX(c, p, NJ - 1, q); }
                const double* S = (b < 2) ? &c->sI[f] : &c->sJ[f];
                st5 W = prim(ld(c, U, node));
                v3 u = mk3(W.v[1], W.v[2], W.v[3]);
                double pt = f_pt(c, W.v[0], u, W.v[4]);
                double Tt = f_Tt(c, W.v[0], u, W.v[4]);
                sp += S[0] * (W.v[0] * u.x * pt) + S[1] * (W.v[0] * u.y * pt) + S[2] * (W.v[0] * u.z * pt);
                stt += S[0] * (W.v[0] * u.x * Tt) + S[1] * (W.v[0] * u.y * Tt) + S[2] * (W.v[0] * u.z * Tt);
                if (c->p.fluid_real) {
                    double ht = f_ht(c, W.v[0], u, W.v[4]);
                    double sv = f_s(c, W.v[0], u, W.v[4]);
                    sht += S[0] * (W.v[0] * u.x * ht) + S[1] * (W.v[0] * u.y * ht) + S[2] * (W.v[0] * u.z * ht);
                    ssv += S[0] * (W.v[0] * u.x * sv) + S[1] * (W.v[0] * u.y * sv) + S[2] * (W.v[0] * u.z * sv);
                }
            }
        double mfl = mf[b];
        if (fabs(mfl) < 1e-12) mfl = (mfl >= 0.0 ? 1e-12 : -1e-12);
        ptm[s] = sp / mfl;
        ttm[s] = stt / mfl;
        htm[s] = sht / mfl;
        sm[s] = ssv / mfl;
    }
    double PR = (fabs(ptm[0]) > 1e-12) ? ptm[1] / ptm[0] : 1.0;
    double TR = (fabs(ttm[0]) > 1e-12) ? ttm[1] / ttm[0] : 1.0;
    out[7] = PR;
    out[8] = TR;
    double eta;
    if (c->p.fluid_real) {
        /* solver.cpp:1769-1776, FluidBase::computeTotalEfficiency_h */
        double ht_in = htm[0], s_in = sm[0], pt_out = ptm[1], ht_out = htm[1];
        double ht_out_s = lut_h_p_s(c, pt_out, s_in);
        double dh = ht_out - ht_in, dhs = ht_out_s - ht_in;
        if (fabs(dh) < 1e-12) eta = 1.0;
        else if (dh >= 0.0) eta = dhs / dh;
        else eta = (fabs(dhs) < 1e-12) ? 1.0 : dh / dhs;
    } else {
        double g = c->p.gamma;
        double dT = TR - 1.0, dTs = pow(PR, (g - 1.0) / g) - 1.0;
        if (fabs(dT) < 1e-12) eta = 1.0;
        else if (dT >= 0.0) eta = dTs / dT;
        else eta = (fabs(dTs) < 1e-12) ? 1.0 : dT / dTs;
    }
    out[9] = eta;
}

/* --------------------------------------------------------------------- set-up */
static void build_maps(tbo_case* c) {
    const long NI = c->p.ni, NJ = c->p.nj, NK = c->p.nk;
    c->mapI = (int*)malloc(sizeof(int) * 2 * NJ * NK);
    c->mapJ = (int*)malloc(sizeof(int) * NI * 2 * NK);
    c->mapK = (int*)malloc(sizeof(int) * NI * NJ * 2);
    for (long n = 0; n < 2 * NJ * NK; ++n) c->mapI[n] = -1;
    for (long n = 0; n < NI * 2 * NK; ++n) c->mapJ[n] = -1;
    for (long n = 0; n < NI * NJ * 2; ++n) c->mapK[n] = -1;
    /* solver.cpp:491-561 */
    for (int q = 0; q < c->npatches; ++q) {
        const tbo_patch* pa = &c->patches[q];
        if (pa->i0 == pa->i1) {
            long t = (pa->i0 == 0) ? 0 : 1;
            for (long j = pa->j0; j < pa->j1; ++j)
                for (long k = pa->k0; k < pa->k1; ++k) c->mapI[(t * NJ + j) * NK + k] = q;
        } else if (pa->j0 == pa->j1) {
            long t = (pa->j0 == 0) ? 0 : 1;
            for (long i = pa->i0; i < pa->i1; ++i)
                for (long k = pa->k0; k < pa->k1; ++k) c->mapJ[(i * 2 + t) * NK + k] = q;
        } else if (pa->k0 == pa->k1) {
            long t = (pa->k0 == 0) ? 0 : 1;
            for (long i = pa->i0; i < pa->i1; ++i)
                for (long j = pa->j0; j < pa->j1; ++j) c->mapK[(i * NJ + j) * 2 + t] = q;
        }
    }
}

tbo_case* tbo_create(const tbo_params* p, const double* x, const double* y, const double* z,
                     const tbo_fields* f, const tbo_patch* patches, int npatches, char* err, int errlen) {
    tbo_case* c = (tbo_case*)calloc(1, sizeof(tbo_case));
    c->p = *p;
    c->N = p->ni * p->nj * p->nk;
    long N = c->N;
    c->x = dalloc(N); c->y = dalloc(N); c->z = dalloc(N);
    memcpy(c->x, x, N * 8); memcpy(c->y, y, N * 8); memcpy(c->z, z, N * 8);
#define CPF(dst, src) do { if (f && f->src) { c->dst = dalloc(N); memcpy(c->dst, f->src, N * 8); } } while (0)
    CPF(blockage, blockage); CPF(nax, normal_axial); CPF(nrad, normal_radial); CPF(ntan, normal_tangential);
    CPF(rpm, rpm); CPF(present, blade_present); CPF(nblades, number_blades); CPF(metal, metal_angle);
    CPF(slen, streamwise_length); CPF(beta0, lift_drag_beta0); CPF(kp, lift_drag_kp);
    CPF(pivot, deviation_angle_pivot); CPF(stall, deviation_angle_stall);
#undef CPF
    c->patches = (tbo_patch*)malloc(sizeof(tbo_patch) * (npatches > 0 ? npatches : 1));
    memcpy(c->patches, patches, sizeof(tbo_patch) * npatches);
    c->npatches = npatches;
    for (int q = 0; q < npatches; ++q) {
        if (patches[q].type == TBO_BC_PERIODIC && patches[q].values[0] == 0) {
            c->periodic = 1;
            c->per_angle_deg = patches[q].values[2];
            if (c->per_angle_deg == 360.0) c->per_angle_deg = 0.0;
            c->per_angle_rad = c->per_angle_deg * M_PI / 180.0;
        }
    }
    build_geometry(c);
    build_maps(c);
    /* check every boundary face has a patch (solver.cpp:563-588) */
    const long NI = p->ni, NJ = p->nj, NK = p->nk;
    int bad = 0;
    for (long n = 0; n < 2 * NJ * NK; ++n) bad |= c->mapI[n] < 0;
    if (p->topology != TBO_TOPO_1D) for (long n = 0; n < NI * 2 * NK; ++n) bad |= c->mapJ[n] < 0;
    if (p->topology == TBO_TOPO_3D) for (long n = 0; n < NI * NJ * 2; ++n) bad |= c->mapK[n] < 0;
    if (bad) {
        if (err) snprintf(err, errlen, "Missing boundary condition on some boundary face");
        tbo_destroy(c);
        return NULL;
    }
    /* outlet pressure profiles: Solver::buildBoundaryDataStructures (solver.cpp:296-340) */
    for (int q = 0; q < npatches; ++q) {
        c->patch_prof[q] = -1;
        const tbo_patch* pa = &patches[q];
        if (pa->type != TBO_BC_RADIAL_EQ && pa->type != TBO_BC_THROTTLE && pa->type != TBO_BC_MASS_FLOW) continue;
        int radial = 1;
        if (pa->type == TBO_BC_MASS_FLOW) {
            int can = (pa->i0 == pa->i1 && pa->i0 != 0 && pa->j1 > pa->j0);
            if (pa->nvalues >= 4) radial = (pa->values[3] > 0.5) && can;
            else radial = can && (p->topology == TBO_TOPO_AXI || p->bfm_active);
        }
        int np = c->nprof++;
        c->prof_patch[np] = q;
        c->patch_prof[q] = np;
        c->prof_radial[np] = radial;
        long n = (pa->j1 > pa->j0) ? pa->j1 - pa->j0 : 1;
        c->prof_n[np] = n;
        c->prof_p[np] = dalloc(n);
        c->prof_r[np] = dalloc(n);
        if (radial)
            for (long j = pa->j0; j < pa->j1; ++j)
                c->prof_r[np][j - pa->j0] = c->radius[IDX(c, pa->i1 - 1, j, 0)];
    }
    c->U0 = dalloc(5 * N); c->U = dalloc(5 * N); c->R = dalloc(5 * N); c->dt = dalloc(N);
    c->finv = dalloc(3 * N); c->fvisc = dalloc(3 * N); c->dev = dalloc(N);
    /* no-slip wall nodes and velocities, patches in file order (solver.cpp:369-389,
     * boundary_base.cpp:227-268); fewer than three values leave the wall at rest */
    for (int q = 0; q < npatches; ++q) {
        const tbo_patch* pa = &patches[q];
        if (pa->type != TBO_BC_NO_SLIP) continue;
        if (!c->nsw) { c->nsw = (int*)calloc((size_t)N, sizeof(int)); c->wallv = dalloc(3 * N); }
        long r0[3] = {pa->i0, pa->j0, pa->k0}, r1[3] = {pa->i1, pa->j1, pa->k1}, nn[3] = {NI, NJ, NK};
        for (int d = 0; d < 3; ++d)
            if (r0[d] == r1[d]) {
                if (r0[d] == 0) r1[d] = 1;
                else { r0[d] = nn[d] - 1; r1[d] = nn[d]; }
                break;
            }
        double wv[3] = {0.0, 0.0, 0.0};
        if (pa->nvalues >= 3) { wv[0] = pa->values[0]; wv[1] = pa->values[1]; wv[2] = pa->values[2]; }
        for (long i = r0[0]; i < r1[0]; ++i)
            for (long j = r0[1]; j < r1[1]; ++j)
                for (long k = r0[2]; k < r1[2]; ++k) {
                    long n = IDX(c, i, j, k);
                    c->nsw[n] = 1;
                    c->wallv[3 * n] = wv[0]; c->wallv[3 * n + 1] = wv[1]; c->wallv[3 * n + 2] = wv[2];
                }
    }
    if (p->viscous) {
        if (!c->nsw) { c->nsw = (int*)calloc((size_t)N, sizeof(int)); c->wallv = dalloc(3 * N); }
        c->grad = dalloc(12 * N);
    }
    return c;
}

void tbo_destroy(tbo_case* c) {
    if (c) for (int t = 0; t < TBO_LUT_COUNT; ++t) free(c->lutdata[t]);
    if (!c) return;
    double* ptrs[] = {c->x, c->y, c->z, c->dual, c->sI, c->sJ, c->sK, c->cI, c->cJ, c->cK, c->vol,
                      c->radius, c->theta, c->gradb, c->blockage, c->nax, c->nrad, c->ntan, c->rpm,
                      c->present, c->nblades, c->metal, c->slen, c->beta0, c->kp, c->U0, c->U, c->R,
                      c->dt, c->finv, c->fvisc, c->dev, c->pivot, c->stall, c->wallv, c->grad};
    for (size_t q = 0; q < sizeof(ptrs) / sizeof(ptrs[0]); ++q) free(ptrs[q]);
    for (int q = 0; q < c->nprof; ++q) { free(c->prof_p[q]); free(c->prof_r[q]); }
    free(c->mapI); free(c->mapJ); free(c->mapK); free(c->patches); free(c->nsw);
    free(c);
}

long tbo_num_nodes(const tbo_case* c) { return c->N; }

static void init_profiles(tbo_case* c) {
    /* solver.cpp:802-816: the Greitzer plenum starts at the outlet hub node's static
     * pressure and temperature with the turbomachinery mass flow in and out */
    if (c->p.greitzer) {
        double t[10];
        tbo_turbo(c, c->U0, t);
        st5 W = prim(ld(c, c->U0, IDX(c, c->p.ni - 1, 0, 0)));
        v3 u = mk3(W.v[1], W.v[2], W.v[3]);
        c->gz_p = f_p(c, W.v[0], u, W.v[4]);
        c->gz_T = f_T(c, W.v[0], u, W.v[4]);
        c->gz_mout = t[6];
    }
    /* solver.cpp:818-834: profile = static pressure of the current solution at i_min-1,
     * k = 0; a mass-flow patch with an initial pressure (second value > 0) takes that */
    for (int q = 0; q < c->nprof; ++q) {
        const tbo_patch* pa = &c->patches[c->prof_patch[q]];
        if (pa->type == TBO_BC_MASS_FLOW && pa->nvalues >= 2 && pa->values[1] > 0.0) {
            for (long m = 0; m < c->prof_n[q]; ++m) c->prof_p[q][m] = pa->values[1];
            continue;
        }
        long i = pa->i0 > 0 ? pa->i0 - 1 : 0;
        long k = 0;
        for (long j = pa->j0; j < pa->j1 && k < c->prof_n[q]; j++) {
            st5 W = prim(ld(c, c->U0, IDX(c, i, j, 0)));
            c->prof_p[q][k] = f_p(c, W.v[0], mk3(W.v[1], W.v[2], W.v[3]), W.v[4]);
            k++;
        }
    }
}

/* solver.cpp:853-888 + fluid_ideal.cpp:155-170 */
void tbo_init_state(tbo_case* c) {
    const double g = c->p.gamma, Rg = c->p.r_gas;
    v3 dir = mk3(c->p.init_dir[0], c->p.init_dir[1], c->p.init_dir[2]);
    double rho, et;
    v3 vel;
    if (c->p.fluid_real) {
        /* FluidReal::computeInitFields (src/fluid_real.cpp:375-393) */
        const double p0 = c->p.init_pressure, T0 = c->p.init_temperature;
        rho = lut_interp(c, TBO_LUT_RHO_PT, p0, T0);
        double e = lut_interp(c, TBO_LUT_E_PT, p0, T0);
        double a = lut_interp(c, TBO_LUT_A, rho, e);
        double dm = mag3(dir);
        v3 dn = (dm > 1e-12) ? div3(dir, dm) : mk3(1.0, 0.0, 0.0);
        vel = mul3(mul3(dn, a), c->p.init_mach);
        et = e + 0.5 * dot3(vel, vel);
    } else {
        double a = sqrt(g * Rg * c->p.init_temperature);
        vel = mul3(mul3(div3(dir, mag3(dir)), a), c->p.init_mach);
        rho = c->p.init_pressure / (Rg * c->p.init_temperature);
        double e = c->p.init_pressure / ((g - 1) * rho);
        double vm = mag3(vel);
        et = e + 0.5 * (vm * vm);
    }
    for (long n = 0; n < c->N; ++n) {
        c->U0[n] = rho;
        c->U0[c->N + n] = rho * vel.x;
        c->U0[2 * c->N + n] = rho * vel.y;
        c->U0[3 * c->N + n] = rho * vel.z;
        c->U0[4 * c->N + n] = rho * et;
    }
    init_profiles(c);
    if (c->p.viscous) wall_momentum(c, c->U0);   /* preprocessSolution before the loop, solver.cpp:1324 */
}

void tbo_set_state(tbo_case* c, const double* U, int reset_profiles) {
    memcpy(c->U0, U, sizeof(double) * 5 * c->N);
    if (reset_profiles) init_profiles(c);
    if (c->p.viscous) wall_momentum(c, c->U0);   /* preprocessSolution before the loop, solver.cpp:1324 */
}
void tbo_get_state(const tbo_case* c, double* U) { memcpy(U, c->U0, sizeof(double) * 5 * c->N); }

/* Outlet pressure profile q (prof_n values); returns the count, -1 if q is out of range. */
long tbo_get_profile(const tbo_case* c, int q, double* out) {
    if (q < 0 || q >= c->nprof) return -1;
    if (out) memcpy(out, c->prof_p[q], sizeof(double) * c->prof_n[q]);
    return c->prof_n[q];
}

/* test hooks: overwrite outlet profile q; the boundary flux of patch pidx at face
 * (fi, fj, fk) as the residual computes it (bc_flux) */
long tbo_set_profile(tbo_case* c, int q, const double* in) {
    if (q < 0 || q >= c->nprof) return -1;
    memcpy(c->prof_p[q], in, sizeof(double) * c->prof_n[q]);
    return c->prof_n[q];
}
void tbo_patch_flux(const tbo_case* c, int pidx, const double* Ui, const double* S, const long* face, long it,
                    double* out) {
    st5 U;
    memcpy(U.v, Ui, 40);
    st5 F = bc_flux(c, pidx, U, mk3(S[0], S[1], S[2]), face[0], face[1], face[2], it);
    memcpy(out, F.v, 40);
}

void tbo_side_outputs(const tbo_case* c, double* finv, double* fvisc, double* dev, double* dt) {
    if (finv) memcpy(finv, c->finv, 24 * c->N);
    if (fvisc) memcpy(fvisc, c->fvisc, 24 * c->N);
    if (dev) memcpy(dev, c->dev, 8 * c->N);
    if (dt) memcpy(dt, c->dt, 8 * c->N);
}

void tbo_residual(tbo_case* c, const double* U, long it, int update_profiles, double* R) {
    if (update_profiles) update_radial_profiles(c, U);
    compute_residuals(c, U, R, it);
}

/* solver.cpp:2790-2897 alone: R = minus the geometric, Gong and body-force sources of U
 * (test hook for the closed-form checks of those terms) */
void tbo_source_residual(tbo_case* c, const double* U, int which, double* R) {
    memset(R, 0, sizeof(double) * 5 * c->N);
    source_terms(c, U, R, which);
}

void tbo_set_time(tbo_case* c, double t) { c->t_phys = t; }
double tbo_get_time(const tbo_case* c) { return c->t_phys; }

/* GreitzerModel::initializeState + computePlenumPressure (greitzer_model.cpp:16-56) for a
 * sequence of compressor mass flows; pout / mout receive the plenum pressure and outlet
 * mass flow after each step (test hook of the plenum model alone) */
void tbo_greitzer_steps(tbo_case* c, double p0, double mout0, double T0, const double* mflow, int n,
                        double* pout, double* mout) {
    c->gz_p = p0;
    c->gz_mout = mout0;
    c->gz_T = T0;
    for (int q = 0; q < n; ++q) {
        pout[q] = greitzer_plenum_pressure(c, mflow[q]);
        mout[q] = c->gz_mout;
    }
}

/* solver.cpp:1298-1424 (explicit RK branch) + 1426-1439 + 2985-3015 */
long tbo_run(tbo_case* c, long n, double* logres, double* turbo, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const long N = c->N;
    double init[5] = {0, 0, 0, 0, 0};
    int has_init = 0;
    long converged = 0;
    for (long it = 1; it <= n; ++it) {
        if (turbo) {
            double t[10];
            tbo_turbo(c, c->U0, t);
            turbo[(it - 1) * 4 + 0] = t[6]; turbo[(it - 1) * 4 + 1] = t[7];
            turbo[(it - 1) * 4 + 2] = t[8]; turbo[(it - 1) * 4 + 3] = t[9];
        }
        tbo_timestep(c, c->U0, c->dt);
        memcpy(c->U, c->U0, sizeof(double) * 5 * N);
        update_radial_profiles(c, c->U);
        if (c->p.viscous) wall_momentum(c, c->U);   /* preprocessSolution, solver.cpp:3189-3204 */
        for (int s = 0; s < c->p.n_stages; ++s) {
            compute_residuals(c, c->U, c->R, it);
            update_solution(c, c->U0, c->U, c->R, c->p.rk[s]);
            if (c->periodic) periodic_solution(c, c->U);
        }
        memcpy(c->U0, c->U, sizeof(double) * 5 * N);
        /* solver.cpp:1389: the physical time advances by the minimum time step */
        double dtm = c->dt[0];
        for (long q = 1; q < N; ++q) dtm = (c->dt[q] < dtm) ? c->dt[q] : dtm;
        c->t_phys += dtm;
        double lr[5];
        int st = log_norms(c, c->R, lr);
        if (st < 0) return st;
        if (logres) for (int v = 0; v < 5; ++v) logres[(it - 1) * 5 + v] = lr[v];
        if (!has_init) { memcpy(init, lr, sizeof(lr)); has_init = 1; }
        else if (it <= 5) for (int v = 0; v < 5; ++v) init[v] = dmax(init[v], lr[v]);
        if (c->p.steady && !converged) {
            int ok = 1;
            for (int v = 0; v < 5; ++v) {
                if (v == 3 && c->p.topology != TBO_TOPO_3D && c->p.topology != TBO_TOPO_AXI) continue;
                if (v == 2 && c->p.topology == TBO_TOPO_1D) continue;
                if (init[v] <= -15.99) continue;
                if (lr[v] <= -15.0) continue;
                if (lr[v] >= init[v] - (double)c->p.residual_drop) { ok = 0; break; }
            }
            if (ok) { converged = it; return converged; }
        }
    }
    return converged;
}
