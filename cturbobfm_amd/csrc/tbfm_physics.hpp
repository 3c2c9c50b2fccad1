// Point physics on the device: ideal-gas relations, face fluxes, boundary fluxes and
// body-force sources.  Each function names the reference lines it reproduces; the
// operation order is the reference's so that FP64 results are bit-identical.
#pragma once
#include "tbfm_kernels.hpp"

namespace tbfm {
namespace dev {

// ---------------------------------------------------------------- ideal gas
// static energy et - 0.5*pow(|u|,2) with |u| = sqrt(u.u)   (src/fluid_ideal.cpp:70-77)
template <typename T>
__device__ __forceinline__ T static_energy(T u, T v, T w, T et) {
    T m = sqrt(u * u + v * v + w * w);
    return et - T(0.5) * (m * m);
}

// Fluid of a kernel instance: FL = 0 the ideal gas of constant gamma
// (src/fluid_ideal.cpp), FL = 1 the look-up-table fluid (src/fluid_real.cpp), FP64 only.
template <int FL>
struct Gas;
template <>
struct Gas<0> {
    double g, gm1;
    __device__ __forceinline__ explicit Gas(const CaseDev& c) : g(c.gamma), gm1(c.gm1) {}
};
template <>
struct Gas<1> {
    const LutSet* L;
    __device__ __forceinline__ explicit Gas(const CaseDev& c) : L(c.lut) {}
    __device__ __forceinline__ double at(int t, double x, double y) const { return lut_interp(L->t[t], x, y); }
};

// Static pressure of primitive state W.  Ideal: (g-1) rho (et - 0.5 pow(|u|,2))
// (fluid_ideal.cpp:22-24, 70-77); real: p(rho, et - 0.5 u.u) (FluidBase composition,
// fluid_real.cpp:211-213).
template <typename T, int FL>
__device__ __forceinline__ T pressure_w(const Gas<FL>& gs, const T W[5]) {
    if constexpr (FL == 0) {
        T e = static_energy(W[1], W[2], W[3], W[4]);
        return T(gs.gm1) * W[0] * e;
    } else {
        static_assert(sizeof(T) == 8, "the look-up-table fluid runs in FP64");
        double e = W[4] - 0.5 * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]);
        return gs.at(LUT_P, W[0], e);
    }
}

// Speed of sound from rho, u, et (computeSoundSpeed_rho_u_et): ideal sqrt(g p / rho) with
// the pressure p of the same state (fluid_ideal.cpp:83-87); real a(rho, et - 0.5 u.u).
template <int FL>
__device__ __forceinline__ double sound_speed_w(const Gas<FL>& gs, const double W[5], double p) {
    if constexpr (FL == 0) return sqrt(gs.g * p / W[0]);
    else return gs.at(LUT_A, W[0], W[4] - 0.5 * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]));
}

// The same in the source type T of the body force (T = float: the ideal gas only).
template <int FL, typename T>
__device__ __forceinline__ T sound_speed_wt(const Gas<FL>& gs, const T W[5], T p) {
    if constexpr (sizeof(T) == 8) return sound_speed_w<FL>(gs, W, p);
    else return sqrt(T(gs.g) * p / W[0]);
}

// Speed of sound from p and rho (computeSoundSpeed_p_rho): ideal sqrt(g p / rho); real
// a(rho, e(p, rho)) with the bisection for e (fluid_real.cpp:254-257).
template <int FL>
__device__ __forceinline__ double sound_speed_p_rho(const Gas<FL>& gs, double p, double rho) {
    if constexpr (FL == 0) return sqrt(gs.g * p / rho);
    else return gs.at(LUT_A, rho, lut_e_p_rho(*gs.L, p, rho));
}

// W = (rho, rhoU/rho, rhoV/rho, rhoW/rho, rhoE/rho) (src/math_utils.cpp:49-57), the
// static pressure p and the JST spectral factor |u| + a (advection_jst.cpp:40-45).
// Ideal: p = (g-1) rho e, a = sqrt(g p / rho) (fluid_ideal.cpp:22-24, 83-87); real:
// p(rho, e), a(rho, e) with e = et - 0.5 u.u.
template <typename T, int FL>
__device__ __forceinline__ void derive(const T U[5], T W[5], T& p, T& rf, const Gas<FL>& gs) {
    W[0] = U[0];
    W[1] = U[1] / U[0];
    W[2] = U[2] / U[0];
    W[3] = U[3] / U[0];
    W[4] = U[4] / U[0];
    if constexpr (FL == 0) {
        T m = sqrt(W[1] * W[1] + W[2] * W[2] + W[3] * W[3]);
        T e = W[4] - T(0.5) * (m * m);
        p = T(gs.gm1) * W[0] * e;
        T a = sqrt(T(gs.g) * p / W[0]);
        rf = m + a;
    } else {
        static_assert(sizeof(T) == 8, "the look-up-table fluid runs in FP64");
        double q = W[1] * W[1] + W[2] * W[2] + W[3] * W[3];
        double m = sqrt(q);
        double e = W[4] - 0.5 * q;
        p = gs.at(LUT_P, W[0], e);
        rf = m + gs.at(LUT_A, W[0], e);
    }
}

// Physical flux through the unit normal n of primitive state W with static pressure p
// (src/math_utils.cpp:71-86; h_t = e_t + p / rho, fluid_ideal.cpp:111-115).
template <typename T>
__device__ __forceinline__ void phys_flux_pn(const T W[5], T nx, T ny, T nz, T p, T F[5]) {
    T un = W[1] * nx + W[2] * ny + W[3] * nz;
    T ht = W[4] + p / W[0];
    F[0] = W[0] * un;
    F[1] = W[0] * un * W[1] + p * nx;
    F[2] = W[0] * un * W[2] + p * ny;
    F[3] = W[0] * un * W[3] + p * nz;
    F[4] = W[0] * un * ht;
}

// Physical flux through S (n = S/|S|).  Returns |S| so the caller can scale by it, as the
// residual assembly does.
template <typename T, int FL>
__device__ __forceinline__ T phys_flux(const T W[5], T sx, T sy, T sz, const Gas<FL>& gs, T F[5]) {
    T sm = sqrt(sx * sx + sy * sy + sz * sz);
    T nx = sx / sm, ny = sy / sm, nz = sz / sm;
    phys_flux_pn<T>(W, nx, ny, nz, pressure_w<T, FL>(gs, W), F);
    return sm;
}

// Same with the unit normal n = S/|S| supplied (precomputed with identical operations).
template <typename T, int FL>
__device__ __forceinline__ void phys_flux_n(const T W[5], T nx, T ny, T nz, const Gas<FL>& gs, T F[5]) {
    phys_flux_pn<T>(W, nx, ny, nz, pressure_w<T, FL>(gs, W), F);
}

// JST pressure sensor |p1 - 2 p2 + p3| / (p1 + 2 p2 + p3)   (advection_jst.cpp:47-57)
template <typename T>
__device__ __forceinline__ T jst_sensor(T p1, T p2, T p3) {
    return fabs((p1 - 2 * p2 + p3) / (p1 + 2 * p2 + p3));
}

// JST face flux times |S| (advection_jst.cpp:4-38; k2 = 0.5, k4 = 0.02, c4 = 2,
// advection_jst.hpp:23).  Conserved differences are formed in the storage type ST of the
// states (double, or float in the FP32 stage kernel); the flux type FT carries the rest.
template <typename FT, int FL, typename ST, typename RT>
__device__ __forceinline__ void jst_face(const ST Ull[5], const ST Ul[5], const ST Ur[5], const ST Urr[5],
                                         const ST Wl[5], const ST Wr[5], ST rfl, ST rfr, ST sl, ST sr, RT nx,
                                         RT ny, RT nz, RT area, const Gas<FL>& gs, double out[5]) {
    FT Wavg[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) Wavg[v] = (FT(Wl[v]) + FT(Wr[v])) * FT(0.5);
    FT r = FT(smax(rfl, rfr));
    FT s = FT(smax(sl, sr));
    FT psi2 = FT(0.5) * s * r;
    FT t4 = FT(0.02) * r - FT(2.0) * psi2;
    FT psi4 = (FT(0.0) < t4) ? t4 : FT(0.0);
    FT F[5];
    phys_flux_n<FT, FL>(Wavg, FT(nx), FT(ny), FT(nz), gs, F);
    const FT sm = FT(area);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        ST d1 = Ur[v] - Ul[v];
        ST d3 = ((Urr[v] - Ur[v]) - d1 * 2) + (Ul[v] - Ull[v]);
        FT dis = FT(d1) * psi2 - FT(d3) * psi4;
        F[v] -= dis;
        out[v] = double(F[v] * sm);
    }
}

// MUSCL reconstruction of primitives (advection_base.cpp:10-65)
template <typename T>
__device__ __forceinline__ T muscl_limiter(int lim, T s) {
    switch (lim) {
        case 1: return (s <= T(0)) ? T(0) : (s * s + s) / (T(1) + s * s);
        case 2: return (s + fabs(s)) / (T(1) + fabs(s));
        case 3: {
            T m = (T(1) < s) ? T(1) : s;   // std::min(1, s)
            return (T(0) < m) ? m : T(0);  // std::max(0, m)
        }
        default: return T(0);
    }
}
template <typename T>
__device__ __forceinline__ T muscl_ratio(T l, T c, T r) {
    T dR = r - c;
    T dL = c - l;
    return (fabs(dR) < T(1E-12)) ? T(0) : dL / dR;
}

template <typename T>
__device__ __forceinline__ T roe_avg(T rl, T rr, T pl, T pr) {
    return (sqrt(rl) * pl + sqrt(rr) * pr) / (sqrt(rl) + sqrt(rr));
}

// Roe flux with Vinokur-Montagne averaging and entropy fix 0.01, times |S|
// (src/advection_roe.cpp:6-221, include/advection_roe.hpp:63)
template <typename FT, int FL>
__device__ __forceinline__ void roe_face(const double Wll_[5], const double Wl_[5], const double Wr_[5],
                                         const double Wrr_[5], int muscl, int lim, double nx_, double ny_,
                                         double nz_, double area_, const Gas<FL>& gs, double out[5]) {
    FT Wl[5], Wr[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) { Wl[v] = FT(Wl_[v]); Wr[v] = FT(Wr_[v]); }
    if (muscl) {
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            FT sL = muscl_ratio<FT>(FT(Wll_[v]), FT(Wl_[v]), FT(Wr_[v]));
            FT sR = muscl_ratio<FT>(FT(Wl_[v]), FT(Wr_[v]), FT(Wrr_[v]));
            FT lL = muscl_limiter<FT>(lim, sL), lR = muscl_limiter<FT>(lim, sR);
            FT wl = FT(Wl_[v]), wr = FT(Wr_[v]), wrr = FT(Wrr_[v]);
            Wl[v] = wl + lL * (wr - wl) * FT(0.5);
            Wr[v] = wr - lR * (wrr - wr) * FT(0.5);
        }
    }
    // normal triad (:41-58); n1 = S.normalized() is supplied
    const FT sm = FT(area_);
    const FT n1x = FT(nx_), n1y = FT(ny_), n1z = FT(nz_);
    FT t2x, t2y, t2z;
    if (n1x == FT(0) && n1y == FT(0)) { t2x = FT(1); t2y = FT(0); t2z = FT(0); }
    else { t2x = -n1y; t2y = n1x; t2z = FT(0); }
    FT m2 = sqrt(t2x * t2x + t2y * t2y + t2z * t2z);
    FT n2x, n2y, n2z;
    if (m2 == FT(0)) { n2x = n2y = n2z = FT(0); }
    else { n2x = t2x / m2; n2y = t2y / m2; n2z = t2z / m2; }
    FT n3x = n1y * n2z - n1z * n2y, n3y = n1z * n2x - n1x * n2z, n3z = n1x * n2y - n1y * n2x;
    // rotated primitives (:60-64)
    FT L[5], R[5];
    L[0] = Wl[0]; L[4] = Wl[4];
    L[1] = Wl[1] * n1x + Wl[2] * n1y + Wl[3] * n1z;
    L[2] = Wl[1] * n2x + Wl[2] * n2y + Wl[3] * n2z;
    L[3] = Wl[1] * n3x + Wl[2] * n3y + Wl[3] * n3z;
    R[0] = Wr[0]; R[4] = Wr[4];
    R[1] = Wr[1] * n1x + Wr[2] * n1y + Wr[3] * n1z;
    R[2] = Wr[1] * n2x + Wr[2] * n2y + Wr[3] * n2z;
    R[3] = Wr[1] * n3x + Wr[2] * n3y + Wr[3] * n3z;
    // Roe averages (:66-144).  The ideal gas forms h_t from the pressure of
    // FluidIdeal::computeStaticEnergy_u_et (pow(|u|,2)) and p, chi, kappa from e of the
    // explicit formula; for the table fluid both energies are et - 0.5 u.u.
    const FT rhoL = L[0], rhoR = R[0];
    FT eL = L[4] - FT(0.5) * (L[1] * L[1] + L[2] * L[2] + L[3] * L[3]);
    FT eR = R[4] - FT(0.5) * (R[1] * R[1] + R[2] * R[2] + R[3] * R[3]);
    FT pLt, pRt, pL, pR, chiL, kapL, chiR, kapR;
    if constexpr (FL == 0) {
        const FT g = FT(gs.g), gm1 = FT(gs.gm1);
        pLt = gm1 * rhoL * static_energy(L[1], L[2], L[3], L[4]);
        pRt = gm1 * rhoR * static_energy(R[1], R[2], R[3], R[4]);
        pL = gm1 * rhoL * eL;
        pR = gm1 * rhoR * eR;
        chiL = (g - FT(1.0)) * eL; kapL = (g - FT(1.0)) * rhoL;
        chiR = (g - FT(1.0)) * eR; kapR = (g - FT(1.0)) * rhoR;
    } else {
        static_assert(sizeof(FT) == 8, "the look-up-table fluid runs in FP64");
        pLt = pL = gs.at(LUT_P, rhoL, eL);
        pRt = pR = gs.at(LUT_P, rhoR, eR);
        chiL = gs.at(LUT_DPDRHO, rhoL, eL); kapL = gs.at(LUT_DPDE, rhoL, eL);
        chiR = gs.at(LUT_DPDRHO, rhoR, eR); kapR = gs.at(LUT_DPDE, rhoR, eR);
    }
    FT htL = L[4] + pLt / rhoL, htR = R[4] + pRt / rhoR;
    FT rhoA = sqrt(rhoL * rhoR);
    FT u1A = roe_avg(rhoL, rhoR, L[1], R[1]);
    FT u2A = roe_avg(rhoL, rhoR, L[2], R[2]);
    FT u3A = roe_avg(rhoL, rhoR, L[3], R[3]);
    FT htA = roe_avg(rhoL, rhoR, htL, htR);
    FT barChi = FT(0.5) * (chiL + chiR), barKap = FT(0.5) * (kapL + kapR);
    FT barRho = FT(0.5) * (rhoL + rhoR), barE = FT(0.5) * (eL + eR);
    FT dP = pR - pL, dRho = rhoR - rhoL, dE = eR - eL;
    FT dRhoHat = (barRho > FT(1e-12)) ? dRho / barRho : FT(0.0);
    FT dEHat = (fabs(barE) > FT(1e-12)) ? dE / fabs(barE) : FT(0.0);
    FT dNorm2 = dRhoHat * dRhoHat + dEHat * dEHat;
    FT chiT = barChi, kapT = barKap;
    if (dNorm2 > FT(1e-12)) {
        FT res = dP - (barChi * dRho + barKap * dE);
        chiT += (res * dRhoHat) / (dNorm2 * barRho);
        kapT += (res * dEHat) / (dNorm2 * fabs(barE));
    }
    if (barKap > FT(0.0) && kapT <= FT(0.0)) kapT = FT(0.01) * barKap;
    FT kapP = (rhoA > FT(1e-12)) ? kapT / rhoA : FT(0.4);
    FT eA = roe_avg(rhoL, rhoR, eL, eR);
    FT chiP = chiT - kapP * eA;
    FT q2 = u1A * u1A + u2A * u2A + u3A * u3A;
    FT hA = htA - FT(0.5) * q2;
    FT a2 = chiP + kapP * hA;
    if (a2 <= FT(0.0)) {
        FT aL, aR;
        if constexpr (FL == 0) {
            const FT g = FT(gs.g);
            aL = sqrt(g * (g - FT(1.0)) * eL);
            aR = sqrt(g * (g - FT(1.0)) * eR);
        } else {
            aL = gs.at(LUT_A, rhoL, eL);
            aR = gs.at(LUT_A, rhoR, eR);
        }
        a2 = FT(0.5) * (aL * aL + aR * aR);
    }
    FT aA = sqrt(a2);
    // eigen-system (:151-167)
    FT lam[5] = {u1A - aA, u1A, u1A, u1A, u1A + aA};
    FT e5 = htA - (aA * aA) / kapP;
    FT rv[5][5] = {{FT(1.0), u1A - aA, u2A, u3A, htA - aA * u1A},
                   {FT(1.0), u1A, u2A, u3A, e5},
                   {FT(0.0), FT(0.0), FT(1.0), FT(0.0), u2A},
                   {FT(0.0), FT(0.0), FT(0.0), FT(1.0), u3A},
                   {FT(1.0), u1A + aA, u2A, u3A, htA + aA * u1A}};
    // wave strengths (:169-182)
    FT dRho2 = R[0] - L[0], dU1 = R[1] - L[1], dU2 = R[2] - L[2], dU3 = R[3] - L[3];
    FT dP2 = pRt - pLt;
    FT w[5];
    w[0] = FT(1.0) / (FT(2.0) * aA * aA) * (dP2 - rhoA * aA * dU1);
    w[1] = dRho2 - dP2 / (aA * aA);
    w[2] = rhoA * dU2;
    w[3] = rhoA * dU3;
    w[4] = FT(1.0) / (FT(2.0) * aA * aA) * (dP2 + rhoA * aA * dU1);
    // assembly (:184-221): physical fluxes along (1,0,0) in the rotated frame
    // (the pressure of either physical flux is pLt / pRt: the same expression)
    FT FLf[5], FRf[5];
    FT one = FT(1.0), zero = FT(0.0);
    phys_flux_pn<FT>(L, one, zero, zero, pLt, FLf);
    phys_flux_pn<FT>(R, one, zero, zero, pRt, FRf);
    FT Fr[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) Fr[i] = (FLf[i] + FRf[i]) * FT(0.5);
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            FT fe = (fabs(lam[j]) < FT(0.01)) ? FT(0.01) : fabs(lam[j]);
            Fr[i] -= FT(0.5) * w[j] * fe * rv[j][i];
        }
    FT F[5];
    F[0] = Fr[0];
    F[1] = Fr[1] * (n1x * one + n1y * zero + n1z * zero) + Fr[2] * (n2x * one + n2y * zero + n2z * zero) +
           Fr[3] * (n3x * one + n3y * zero + n3z * zero);
    F[2] = Fr[1] * (n1x * zero + n1y * one + n1z * zero) + Fr[2] * (n2x * zero + n2y * one + n2z * zero) +
           Fr[3] * (n3x * zero + n3y * one + n3z * zero);
    F[3] = Fr[1] * (n1x * zero + n1y * zero + n1z * one) + Fr[2] * (n2x * zero + n2y * zero + n2z * one) +
           Fr[3] * (n3x * zero + n3y * zero + n3z * one);
    F[4] = Fr[4];
#pragma unroll
    for (int v = 0; v < 5; ++v) out[v] = double(F[v] * sm);
}

// -------------------------------------------------------------- boundary fluxes
// Subsonic inlet (src/boundary_base.cpp:7-124): the outgoing Riemann invariant of the
// interior state and a bisection on the boundary pressure along the inlet total state,
// isentropic for the ideal gas, along the entropy s(p0, T0) for the table fluid.
template <int FL>
__device__ inline void inlet_flux(const CaseDev& c, const Gas<FL>& gs, const double* tab, const double U[5], double sx,
                                  double sy, double sz, double pt, double Tt, double dx, double dy, double dz,
                                  double F[5]) {
    double W[5], p, rf;
    derive<double, FL>(U, W, p, rf, gs);
    // computeSoundSpeed_rho_u_et
    const double aInt = sound_speed_w<FL>(gs, W, p);
    double sm = sqrt(sx * sx + sy * sy + sz * sz);
    double nx = sx / sm, ny = sy / sm, nz = sz / sm;
    double unInt = W[1] * nx + W[2] * ny + W[3] * nz;
    double Z = W[0] * aInt;
    double Cg = dx * nx + dy * ny + dz * nz;
    if (Cg >= 0.0) Cg = -1.0;
    double Rout = p + Z * unInt;
    double pLow = 0.05 * pt, pHigh = 0.999999 * pt;
    double pB = 0.5 * (pLow + pHigh);
    double Wb[5];
    if constexpr (FL == 0) {
        const double g = c.gamma, Rg = c.rgas;
        double cp = g * Rg / (g - 1.0);
        double htIn = cp * Tt;
        const double ex = (g - 1.0) / g;
        // tree node of the midpoint while inside the patch table, and that of pB
        int qn = 0, qB = -1;
        int it = 0;
        bool done = false;
        if (tab) {
            // the tabulated levels, four per round: the speeds of the node and of its
            // descendants 1-3 levels down (contiguous runs of 1, 2, 4, 8 in heap order)
            // are requested together, then the four bisection steps run from registers
            static_assert(INLET_LEVELS % 4 == 0, "whole rounds of four levels");
            const double* Vt = tab + INLET_NODES;
            for (int round = 0; round < INLET_LEVELS / 4 && !done; ++round) {
                double v0 = gld(Vt, (unsigned)qn);
                double v1[2], v2[4], v3[8];
#pragma unroll
                for (int q = 0; q < 2; ++q) v1[q] = gld(Vt, (unsigned)(2 * qn + 1 + q));
#pragma unroll
                for (int q = 0; q < 4; ++q) v2[q] = gld(Vt, (unsigned)(4 * qn + 3 + q));
#pragma unroll
                for (int q = 0; q < 8; ++q) v3[q] = gld(Vt, (unsigned)(8 * qn + 7 + q));
                int pos = 0;   // position of the current node within its level of the subtree
#pragma unroll
                for (int d = 0; d < 4; ++d) {
                    double VMid = v0;
                    if (d == 1) VMid = pos ? v1[1] : v1[0];
                    if (d == 2) VMid = (pos & 2) ? ((pos & 1) ? v2[3] : v2[2]) : ((pos & 1) ? v2[1] : v2[0]);
                    if (d == 3) {
                        const double a = (pos & 2) ? ((pos & 1) ? v3[3] : v3[2]) : ((pos & 1) ? v3[1] : v3[0]);
                        const double b = (pos & 2) ? ((pos & 1) ? v3[7] : v3[6]) : ((pos & 1) ? v3[5] : v3[4]);
                        VMid = (pos & 4) ? b : a;
                    }
                    const int qd = (qn << d) + ((1 << d) - 1) + pos;   // heap index of the node
                    double pMid = 0.5 * (pLow + pHigh);
                    double res = pMid + Z * Cg * VMid - Rout;
                    ++it;
                    pB = pMid;
                    qB = qd;
                    if (fabs(res) < 1e-4 * pt) { done = true; break; }
                    if (res < 0.0) { pLow = pMid; pos = 2 * pos; } else { pHigh = pMid; pos = 2 * pos + 1; }
                }
                if (!done) qn = (qn << 4) + 15 + pos;
            }
        }
        for (; it < 30 && !done; ++it) {
            double pMid = 0.5 * (pLow + pHigh);
            double TMid = Tt * pow(pMid / pt, ex);
            double hMid = cp * TMid;
            double VMid = sqrt(2.0 * dmaxd(0.0, htIn - hMid));
            double res = pMid + Z * Cg * VMid - Rout;
            pB = pMid;
            qB = -1;
            if (fabs(res) < 1e-4 * pt) break;
            if (res < 0.0) pLow = pMid; else pHigh = pMid;
        }
        double TB, VB;
        if (qB >= 0) {
            TB = gld(tab, (unsigned)qB);
            VB = gld(tab + INLET_NODES, (unsigned)qB);
        } else {
            TB = Tt * pow(pB / pt, ex);
            double hB = cp * TB;
            VB = sqrt(2.0 * dmaxd(0.0, htIn - hB));
        }
        double vx = dx * VB, vy = dy * VB, vz = dz * VB;
        double rhoB = pB / (Rg * TB);
        double eB = pB / ((g - 1) * rhoB);
        double etB = eB + 0.5 * (vx * vx + vy * vy + vz * vz);
        Wb[0] = rhoB; Wb[1] = vx; Wb[2] = vy; Wb[3] = vz; Wb[4] = etB;
    } else {
        // boundary_base.cpp:77-111
        const LutSet& L = *gs.L;
        double sIn = gs.at(LUT_S_PT, pt, Tt);
        double htIn = lut_h_p_s(L, pt, sIn);
        for (int it = 0; it < 30; ++it) {
            double pMid = 0.5 * (pLow + pHigh);
            double hMid = lut_h_p_s(L, pMid, sIn);
            double VMid = sqrt(2.0 * dmaxd(0.0, htIn - hMid));
            double res = pMid + Z * Cg * VMid - Rout;
            if (fabs(res) < 1e-4 * pt) { pB = pMid; break; }
            if (res < 0.0) pLow = pMid; else pHigh = pMid;
            pB = pMid;
        }
        double hB = lut_h_p_s(L, pB, sIn);
        double VB = sqrt(2.0 * dmaxd(0.0, htIn - hB));
        double vx = dx * VB, vy = dy * VB, vz = dz * VB;
        double rhoB = gs.at(LUT_RHO_PS, pB, sIn);
        double eB = gs.at(LUT_E_PS, pB, sIn);
        double etB = eB + 0.5 * (vx * vx + vy * vy + vz * vz);
        Wb[0] = rhoB; Wb[1] = vx; Wb[2] = vy; Wb[3] = vz; Wb[4] = etB;
    }
    phys_flux<double, FL>(Wb, sx, sy, sz, gs, F);
}

// Outlet with static pressure and optional ramp (boundary_base.cpp:181-224,
// config.cpp:430-439); supersonic outflow takes the interior flux.  The subsonic
// boundary state keeps the interior velocity and, for the table fluid, the interior
// entropy (boundary_base.cpp:205-211).
template <int FL>
__device__ inline void outlet_flux(const CaseDev& c, const Gas<FL>& gs, const double U[5], double sx, double sy,
                                   double sz, double pout, long it, double F[5]) {
    double W[5], p, rf;
    derive<double, FL>(U, W, p, rf, gs);
    double a = sound_speed_p_rho<FL>(gs, p, W[0]);
    double m = sqrt(W[1] * W[1] + W[2] * W[2] + W[3] * W[3]);
    if (m >= a) {
        phys_flux<double, FL>(W, sx, sy, sz, gs, F);
        return;
    }
    double ratio = double(it) / c.ramp_iter;
    double pB = (ratio >= 1.0) ? pout : c.init_p + (pout - c.init_p) * ratio;
    double rhoB, eB;
    if constexpr (FL == 0) {
        rhoB = pB * W[0] / p;
        eB = pB / ((c.gamma - 1) * rhoB);
    } else {
        double e_int = W[4] - 0.5 * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]);
        double s_int = gs.at(LUT_S, W[0], e_int);
        rhoB = gs.at(LUT_RHO_PS, pB, s_int);
        eB = gs.at(LUT_E_PS, pB, s_int);
    }
    double etB = eB + 0.5 * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]);
    double Wb[5] = {rhoB, W[1], W[2], W[3], etB};
    phys_flux<double, FL>(Wb, sx, sy, sz, gs, F);
}

// Inviscid wall (src/boundary_inviscid_wall.cpp:5-30)
// Transparent end (boundary_transparent.cpp:5-16): the scheme's own face flux with the
// node's state on all four stencil positions, per unit area; S outward.
template <int FL>
__device__ inline void transparent_flux(const CaseDev& c, const Gas<FL>& gs, const double U[5], double sx, double sy,
                                        double sz, double F[5]) {
    const double sm = sqrt(sx * sx + sy * sy + sz * sz);
    const double nx = sx / sm, ny = sy / sm, nz = sz / sm;
    double W[5], p, rf;
    derive<double, FL>(U, W, p, rf, gs);
    if (c.scheme == 0) {
        const double s = jst_sensor<double>(p, p, p);
        jst_face<double, FL>(U, U, U, U, W, W, rf, rf, s, s, nx, ny, nz, 1.0, gs, F);
    } else {
        roe_face<double, FL>(W, W, W, W, c.muscl, c.limiter, nx, ny, nz, 1.0, gs, F);
    }
}

template <int FL>
__device__ inline void wall_flux(const Gas<FL>& gs, const double U[5], double sx, double sy, double sz,
                                 double F[5]) {
    double W[5], p, rf;
    derive<double, FL>(U, W, p, rf, gs);
    double sm = sqrt(sx * sx + sy * sy + sz * sz);
    double nx = sx / sm, ny = sy / sm, nz = sz / sm;
    double un = W[1] * nx + W[2] * ny + W[3] * nz;
    double a = sound_speed_p_rho<FL>(gs, p, W[0]);
    double pw = p + W[0] * un * a;
    if (pw < 0.01 * p) pw = 0.01 * p;
    F[0] = 0.0; F[1] = pw * nx; F[2] = pw * ny; F[3] = pw * nz; F[4] = 0.0;
}

// ------------------------------------------------------------ body-force sources
// Vector3D arithmetic of the reference (include/types.hpp) in the source type T: double
// on the plain path, float for the FP32-source arithmetic of the mixed-precision path.
template <typename T>
struct V3T {
    T x, y, z;
};
using V3 = V3T<double>;
template <typename T>
__device__ __forceinline__ V3T<T> v3(T x, T y, T z) { return V3T<T>{x, y, z}; }
template <typename T>
__device__ __forceinline__ T vdot(V3T<T> a, V3T<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T>
__device__ __forceinline__ T vmag(V3T<T> a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
template <typename T>
__device__ __forceinline__ V3T<T> vsub(V3T<T> a, V3T<T> b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename T>
__device__ __forceinline__ V3T<T> vmul(V3T<T> a, T s) { return v3(a.x * s, a.y * s, a.z * s); }
template <typename T>
__device__ __forceinline__ V3T<T> vdiv(V3T<T> a, T s) { return v3(a.x / s, a.y / s, a.z / s); }
template <typename T>
__device__ __forceinline__ V3T<T> vnrm(V3T<T> a) {
    T m = vmag(a);
    return (m == T(0)) ? v3(T(0), T(0), T(0)) : vdiv(a, m);
}
template <typename T>
__device__ __forceinline__ V3T<T> cyl_of(V3T<T> v, T cs, T sn) {
    return v3(v.x, v.y * cs + v.z * sn, -v.y * sn + v.z * cs);
}
template <typename T>
__device__ __forceinline__ V3T<T> cart_of(V3T<T> v, T cs, T sn) {
    return v3(v.x, v.y * cs - v.z * sn, v.y * sn + v.z * cs);
}

// Compressibility correction of the Hall-Thollet model
// (SourceBFMThollet::computeCompressibilityCorrection, source_bfm_thollet.cpp:50-71):
// relative Mach number m, 1/sqrt(1-m^2) subsonic, 2/(pi sqrt(m^2-1)) supersonic,
// m within 1e-3 of 1 taken as 0.999, capped at 3.
template <typename T>
__device__ inline T thollet_kmach(T wrel, T a) {
    T m = wrel / a;
    if (fabs(m - T(1.0)) < T(1E-03)) m = T(0.999);
    T kp;
    if (m < T(1.0)) kp = T(1.0) / sqrt(T(1.0) - m * m);
    else kp = T(2.0) / T(M_PI) / sqrt(m * m - T(1.0));
    return (kp < T(3.0)) ? kp : T(3.0);
}

// Blockage + Hall / lift-drag / Hall-Thollet source of one node (source_bfm_base.cpp:25-199,
// source_bfm_hall.cpp:3-34, source_bfm_lift_drag.cpp:3-88, source_bfm_thollet.cpp:3-138),
// in the source type T (every input rounded to T first, the results widened back).
// s = slab index.  keepv: the node is stalled under Hall-Thollet, which leaves the
// viscous-force side output at its previous value (source_bfm_thollet.cpp:79-84).
// omd: the node's rotational speed; pf: the perturbation factor (the body force is
// multiplied by 1 + pf, source_bfm_base.cpp:58).
template <int FL, typename T, typename CT>
__device__ inline void bfm_source(const CT& c, const Gas<FL>& gs, long s, double omd, double pf, const double Wd[5],
                                  double Vd, double S[5], double fi[3], double fv[3], double& dev, bool& keepv) {
    const T W[5] = {T(Wd[0]), T(Wd[1]), T(Wd[2]), T(Wd[3]), T(Wd[4])};
    const T V = T(Vd);
    T blk[5] = {0, 0, 0, 0, 0};
    fi[0] = fi[1] = fi[2] = 0.0;
    fv[0] = fv[1] = fv[2] = 0.0;
    dev = 0.0;
    keepv = false;
    if (c.blockage_active) {
        T p = pressure_w<T, FL>(gs, W);
        T ht = W[4] + p / W[0];
        T b = T(c.s_b[s]);
        T dd = W[1] * T(c.s_gbx[s]) + W[2] * T(c.s_gby[s]) + W[3] * T(c.s_gbz[s]);
        T common = T(-1.0) / b * W[0] * dd;
        blk[0] = common * V;
        blk[1] = common * W[1] * V;
        blk[2] = common * W[2] * V;
        blk[3] = common * W[3] * V;
        blk[4] = common * ht * V;
    }
    T body[5] = {0, 0, 0, 0, 0};
    if ((c.sflag[s] & 4) && (c.bfm_model == 1 || c.bfm_model == 2 || c.bfm_model == 4)) {
        // computeFlowState (source_bfm_base.cpp:104-149)
        const T cs = T(c.s_cs[s]), sn = T(c.s_sn[s]), r = T(c.s_r[s]), om = T(omd);
        V3T<T> vcy = cyl_of(v3(W[1], W[2], W[3]), cs, sn);
        V3T<T> rel = vsub(vcy, v3(T(0), T(0), om * r));
        V3T<T> nc = v3(T(c.s_nax[s]), T(c.s_nrad[s]), T(c.s_ntan[s]));
        // deviation angle (:152-166)
        V3T<T> n = vdiv(nc, vmag(nc));
        V3T<T> sp = v3(-rel.y, rel.x, T(0.0));
        sp = vdiv(sp, vmag(sp));
        V3T<T> vin = vsub(rel, vmul(sp, vdot(rel, sp)));
        T sd0 = vdot(vin, n) / vmag(vin);
        // FP32: a relative flow within ~3e-4 rad of the camber normal rounds this sine
        // past +-1 (FP64 needs ~1e-8 rad); the FP32 source keeps it in asin's domain (a
        // NaN from a vanishing relative velocity passes through, as in FP64), the FP64
        // one is the reference's expression as it stands
        if constexpr (sizeof(T) == 4) sd0 = (sd0 < T(-1)) ? T(-1) : ((sd0 > T(1)) ? T(1) : sd0);
        T dv = -asin(sd0);
        // inviscid force direction (:168-199): the camber normal itself once the
        // deviation exceeds the stall deviation (STALLED_BFM_ACTIVE, :178-184)
        const bool stalled = c.stall && dv > T(c.s_stall[s]);
        V3T<T> rd = vnrm(rel), nd = vnrm(nc);
        V3T<T> ver = stalled ? nd : vnrm(vsub(nd, vmul(rd, vdot(nd, rd))));
        V3T<T> sd = vnrm(v3(rel.y, -rel.x, T(0.0)));
        V3T<T> idir = vnrm(vsub(ver, vmul(sd, vdot(ver, sd))));
        dev = dv;
        if (c.bfm_model == 1) {
            T fm = vdot(rel, rel) * T(M_PI) * dv / T(c.s_pitch[s]) / fabs(T(c.s_ntan[s]));
            V3T<T> fcy = vmul(idir, fm);
            V3T<T> fca = cart_of(fcy, cs, sn);
            fi[0] = fca.x; fi[1] = fca.y; fi[2] = fca.z;
            T sv[5] = {T(0), fca.x, fca.y, fca.z, fcy.z * om * r};
#pragma unroll
            for (int v = 0; v < 5; ++v) body[v] = sv[v] * V * W[0];
        } else if (c.bfm_model == 4) {
            // inviscid component (source_bfm_thollet.cpp:23-48)
            const T kmach = thollet_kmach(vmag(rel), sound_speed_wt<FL, T>(gs, W, pressure_w<T, FL>(gs, W)));
            const T rr = vdot(rel, rel);
            const T pitch = T(c.s_pitch[s]), b = T(c.s_b[s]), antan = fabs(T(c.s_ntan[s]));
            T fm = T(c.th_kn) * kmach * rr * T(M_PI) * dv / pitch / antan / b;
            V3T<T> fcy = vmul(idir, fm);
            V3T<T> fca = cart_of(fcy, cs, sn);
            fi[0] = fca.x; fi[1] = fca.y; fi[2] = fca.z;
            T si[5] = {T(0), fca.x, fca.y, fca.z, fcy.z * om * r};
            T sv[5] = {0, 0, 0, 0, 0};
            if (stalled) {
                keepv = true;
            } else {
                // viscous component (:73-138): flat-plate skin friction on the streamwise
                // length (leading-edge fallback applied on the host), plus the off-design
                // loss about the pivot deviation
                T rex = vmag(rel) * T(c.s_stwl[s]) / T(c.th_nu);
                T cf = T(0.0592) * pow(rex, T(-0.2));
                T coef;
                if (!c.th_off_design) coef = T(c.th_kf) * cf;
                else coef = T(c.th_kf) * cf + T(M_PI) * kmach * pow(fabs(dv - T(c.s_piv[s])), T(2.0) * T(c.th_kd));
                T fm2 = rr / (pitch * b * antan) * coef;
                V3T<T> vdir = v3(-vnrm(rel).x, -vnrm(rel).y, -vnrm(rel).z);
                V3T<T> vcyl = vmul(vdir, fm2);
                V3T<T> vcar = cart_of(vcyl, cs, sn);
                fv[0] = vcar.x; fv[1] = vcar.y; fv[2] = vcar.z;
                sv[1] = vcar.x; sv[2] = vcar.y; sv[3] = vcar.z; sv[4] = vcyl.z * om * r;
            }
#pragma unroll
            for (int v = 0; v < 5; ++v) body[v] = si[v] * V * W[0] + sv[v] * V * W[0];
        } else {
            T vmer = sqrt(vcy.x * vcy.x + vcy.y * vcy.y);
            T flow_angle = atan2(rel.z, vmer);
            V3T<T> relc = cart_of(rel, cs, sn);
            T w = vmag(relc);
            T beta0 = T(c.s_beta0[s]);
            T sign = (c.s_ntan[s] >= 0.0) ? T(1.0) : T(-1.0);
            T fm = T(0.5) * w * w * T(2.0) * T(M_PI) / T(c.s_spitch[s]) * sign * (-flow_angle + beta0);
            V3T<T> fcy = vmul(idir, fm);
            V3T<T> fca = cart_of(fcy, cs, sn);
            fi[0] = fca.x; fi[1] = fca.y; fi[2] = fca.z;
            T si[5] = {T(0), fca.x, fca.y, fca.z, fcy.z * om * r};
            T dd = flow_angle - beta0;
            T Kp = T(c.s_kp[s]) + T(2.0) * T(M_PI) * T(c.s_sol[s]) * (dd * dd);
            T fm2 = Kp * w * w / T(c.s_spitch[s]);
            V3T<T> vdir = v3(-vnrm(rel).x, -vnrm(rel).y, -vnrm(rel).z);
            V3T<T> vcyl = vmul(vdir, fm2);
            V3T<T> vcar = cart_of(vcyl, cs, sn);
            fv[0] = vcar.x; fv[1] = vcar.y; fv[2] = vcar.z;
            T sv[5] = {T(0), vcar.x, vcar.y, vcar.z, vcyl.z * om * r};
#pragma unroll
            for (int v = 0; v < 5; ++v) body[v] = si[v] * V * W[0] + sv[v] * V * W[0];
        }
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) S[v] = double(blk[v] + body[v] * (T(1.0) + T(pf)));
}

}  // namespace dev
}  // namespace tbfm
