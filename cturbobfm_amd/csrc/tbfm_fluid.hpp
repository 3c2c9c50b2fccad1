// Look-up-table fluid (FLUID_MODEL = real): the fifteen Table2D tables of a FluidReal and
// the relations built on them, callable from host and device code alike.
//
//   table layout and interpolation   include/fluid_real.hpp:11-60 (Table2D::init,
//                                    Table2D::interpolate)
//   table order                      src/fluid_real.cpp:12-209 (RHO_E: p T a s mu kappa
//                                    dp/drho dp/de; P_T: rho e a s; P_S: rho T e)
//   e(p, rho) bisection              src/fluid_real.cpp:227-252
//   total pressure / temperature     src/fluid_real.cpp:311-341
//   h(p, s) = e(p, s) + p / rho(p, s) FluidBase::computeEnthalpy_p_s
//
// The operation order is the reference's, so FP64 results are bit-identical to it.
#pragma once
#include <hip/hip_runtime.h>

namespace tbfm {

enum LutId {
    LUT_P = 0, LUT_T, LUT_A, LUT_S, LUT_MU, LUT_KAPPA, LUT_DPDRHO, LUT_DPDE,   // f(rho, e)
    LUT_RHO_PT, LUT_E_PT, LUT_A_PT, LUT_S_PT,                                  // f(p, T)
    LUT_RHO_PS, LUT_T_PS, LUT_E_PS,                                            // f(p, s)
    LUT_COUNT
};

// One table: nx * ny values, row-major in x (value (i, j) at d[i * ny + j]); dx, dy are
// the node spacings Table2D::init derives from the bounds.
struct LutTab {
    const double* d;
    int nx, ny;
    double xmin, xmax, ymin, ymax, dx, dy;
};

// The tables of one fluid; d points at host or device memory depending on the copy.
struct LutSet {
    LutTab t[LUT_COUNT];
};

// Table2D::interpolate: clamp into the bounds, cell index by truncation (the last cell
// for the upper bound), bilinear weights, the four terms summed left to right.
__host__ __device__ inline double lut_interp(const LutTab& T, double x, double y) {
    const double xc = (x < T.xmin) ? T.xmin : ((T.xmax < x) ? T.xmax : x);
    const double yc = (y < T.ymin) ? T.ymin : ((T.ymax < y) ? T.ymax : y);
    const double rx = (xc - T.xmin) / T.dx;
    const double ry = (yc - T.ymin) / T.dy;
    // rx, ry lie in [0, n-1] after the clamp; any other value (NaN) is caught below
    unsigned i = (unsigned)rx, j = (unsigned)ry;
    if (i >= (unsigned)(T.nx - 1)) i = (T.nx >= 2) ? (unsigned)(T.nx - 2) : 0u;
    if (j >= (unsigned)(T.ny - 1)) j = (T.ny >= 2) ? (unsigned)(T.ny - 2) : 0u;
    const double wx = rx - (double)i, wy = ry - (double)j;
    const double* r0 = T.d + (long)i * T.ny + j;
    const double* r1 = r0 + T.ny;
    const double f00 = r0[0], f10 = r1[0], f01 = r0[1], f11 = r1[1];
    return (1.0 - wx) * (1.0 - wy) * f00 + wx * (1.0 - wy) * f10 + (1.0 - wx) * wy * f01 + wx * wy * f11;
}

__host__ __device__ inline double lut_at(const LutSet& L, int t, double x, double y) { return lut_interp(L.t[t], x, y); }

// FluidReal::computePressure_rho_T (fluid_real.cpp:262-277): 25 halvings of the T
// table's energy range for T(rho, e) = T, then p(rho, e) at the midpoint.
__host__ __device__ inline double lut_p_rho_T(const LutSet& L, double rho, double T) {
    double e_low = L.t[LUT_T].ymin, e_high = L.t[LUT_T].ymax;
    for (int it = 0; it < 25; ++it) {
        const double e_mid = 0.5 * (e_low + e_high);
        const double T_mid = lut_interp(L.t[LUT_T], rho, e_mid);
        if (T_mid < T) e_low = e_mid;
        else e_high = e_mid;
    }
    return lut_interp(L.t[LUT_P], rho, 0.5 * (e_low + e_high));
}

// FluidReal::computeStaticEnergy_p_rho: bisection on e over the RHO_E energy range,
// at most 30 halvings, tolerance 1e-4 p.
__host__ __device__ inline double lut_e_p_rho(const LutSet& L, double p, double rho) {
    double e_low = L.t[LUT_P].ymin, e_high = L.t[LUT_P].ymax;
    const double p_low = lut_at(L, LUT_P, rho, e_low);
    const double p_high = lut_at(L, LUT_P, rho, e_high);
    if (p <= p_low) return e_low;
    if (p >= p_high) return e_high;
    for (int it = 0; it < 30; ++it) {
        const double e_mid = 0.5 * (e_low + e_high);
        const double p_mid = lut_at(L, LUT_P, rho, e_mid);
        if (fabs(p_mid - p) <= 1e-4 * p) return e_mid;
        if (p_mid < p) e_low = e_mid; else e_high = e_mid;
    }
    return 0.5 * (e_low + e_high);
}

// FluidBase::computeEnthalpy_p_s
__host__ __device__ inline double lut_h_p_s(const LutSet& L, double p, double s) {
    return lut_at(L, LUT_E_PS, p, s) + p / lut_at(L, LUT_RHO_PS, p, s);
}

// FluidReal::computeTotalPressure_rho_u_et: 25 halvings of [p, P_S p_max] for the
// pressure whose enthalpy at the static entropy equals the total enthalpy.  q = u.u.
__host__ __device__ inline double lut_total_pressure(const LutSet& L, double rho, double q, double et) {
    const double e = et - 0.5 * q;
    const double p = lut_at(L, LUT_P, rho, e);
    const double s = lut_at(L, LUT_S, rho, e);
    const double ht = et + p / rho;
    double p_low = p, p_high = L.t[LUT_RHO_PS].xmax;
    for (int it = 0; it < 25; ++it) {
        const double p_mid = 0.5 * (p_low + p_high);
        const double rho_mid = lut_at(L, LUT_RHO_PS, p_mid, s);
        const double e_mid = lut_at(L, LUT_E_PS, p_mid, s);
        const double h_mid = e_mid + p_mid / rho_mid;
        if (h_mid < ht) p_low = p_mid; else p_high = p_mid;
    }
    return 0.5 * (p_low + p_high);
}

// FluidReal::computeTotalTemperature_rho_u_et: T(p_t, s)
__host__ __device__ inline double lut_total_temperature(const LutSet& L, double rho, double q, double et) {
    const double pt = lut_total_pressure(L, rho, q, et);
    const double e = et - 0.5 * q;
    const double s = lut_at(L, LUT_S, rho, e);
    return lut_at(L, LUT_T_PS, pt, s);
}

}  // namespace tbfm
