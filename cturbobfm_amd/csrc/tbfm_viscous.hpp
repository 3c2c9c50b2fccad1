// Laminar viscous terms of the residual (VISCOSITY_ACTIVE) and the no-slip walls, ideal
// gas, FP64: the per-stage velocity and temperature gradients, the viscous face fluxes,
// their per-node accumulation into the stage residual, and the wall momentum reset of the
// first stage's input.  The stage kernel adds the face fluxes in the reference's order
// (after all advection faces, I then J then K, face i before face i+1), so no reduction
// order changes.
#pragma once
#include "tbfm_kernels.hpp"
#include "tbfm_physics.hpp"

namespace tbfm {
namespace dev {

// FluidIdeal::computeMolecularDynamicViscosity (fluid_ideal.cpp:224-231)
__device__ inline double visc_mu(const CaseDev& c, double T) {
    if (c.mu_const) return c.mu_c;
    return c.mu_ref * pow(T / c.mu_tref, 1.5) * ((c.mu_tref + c.mu_s) / (T + c.mu_s));
}

// BoundaryBase::computeStagnationWallFlux (boundary_base.cpp:126-180), ideal gas, per unit
// area: the wall's stagnation pressure, no mass or energy flux
__device__ inline void stagnation_wall_flux(const CaseDev& c, double Sx, double Sy, double Sz, double pt, double Tt,
                                            double ux, double uy, double uz, double F[5]) {
    const double um = sqrt(ux * ux + uy * uy + uz * uz);
    double pb = pt;
    if (um > 0.0) {
        const double g = c.gamma, cp = g * c.rgas / (g - 1.0);
        double Tb = Tt - 0.5 * um * um / cp;
        if (Tb < 0.01 * Tt) Tb = 0.01 * Tt;
        pb = pt * pow(Tb / Tt, g / (g - 1.0));
    }
    const double m = sqrt(Sx * Sx + Sy * Sy + Sz * Sz);
    F[0] = 0.0; F[1] = pb * (Sx / m); F[2] = pb * (Sy / m); F[3] = pb * (Sz / m); F[4] = 0.0;
}

// velocity and temperature of a node (solver.cpp:3039-3044, fluid_ideal.cpp:198-207)
__device__ __forceinline__ void visc_fields(const CaseDev& c, const double* __restrict__ U, long n, double f[4]) {
    const long N = c.N;
    const double rho = U[n];
    const double ux = U[N + n] / rho, uy = U[2 * N + n] / rho, uz = U[3 * N + n] / rho, et = U[4 * N + n] / rho;
    f[0] = ux; f[1] = uy; f[2] = uz;
    f[3] = (et - (ux * ux + uy * uy + uz * uz) * 0.5) * ((c.gamma - 1.0) / c.rgas);
}

// Wall momentum of the iteration's first stage input: U[2] = U[0] with rho * wall velocity
// on the no-slip nodes (preprocessSolution, solver.cpp:3189-3204).  Work: one entry per
// 256 nodes of a viscous case with a moving wall.
__global__ void __launch_bounds__(256) k_wallpre(Launch La) {
    const Work w = La.work[blockIdx.x];
    if (w.cs < 0) return;   // padding entry of an XCD-ordered table
    const CaseDev& c = La.cases[w.cs];
    if (*c.stop) return;
    const long n = (long)w.x * 256 + threadIdx.x, N = c.N;
    if (n >= N) return;
    const double* A = c.U[0];
    double* B = c.U[2];
    const double rho = A[n];
    B[n] = rho;
    B[4 * N + n] = A[4 * N + n];
    const bool wall = c.nsw[n] != 0;
#pragma unroll
    for (int e = 1; e <= 3; ++e) B[e * N + n] = wall ? rho * c.wallv[(e - 1) * N + n] : A[e * N + n];
}

// computeGradientsGreenGauss4 (math_utils.cpp:237-392) of ux, uy, uz, T of the stage
// input: face values interpolated with the host-computed distance weights, the node's
// own value on boundary faces, summed over W E N S B T, divided by the volume.
// Work: one entry per 256 nodes of a viscous case.
__global__ void __launch_bounds__(256) k_vgrad(Launch La) {
    const Work w = La.work[blockIdx.x];
    if (w.cs < 0) return;   // padding entry of an XCD-ordered table
    const CaseDev& c = La.cases[w.cs];
    if (*c.stop) return;
    const long n = (long)w.x * 256 + threadIdx.x, N = c.N;
    if (n >= N) return;
    const long NI = c.ni, NJ = c.nj, NK = c.nk;
    const long i = n % NI, j = (n / NI) % NJ, k = n / (NI * NJ);
    const double* __restrict__ U = stage_input(c, La);
    double fc[4];
    visc_fields(c, U, n, fc);
    const long nb[6] = {i > 0 ? n - 1 : -1, i < NI - 1 ? n + 1 : -1, j < NJ - 1 ? n + NI : -1,
                        j > 0 ? n - NI : -1, k > 0 ? n - NI * NJ : -1, k < NK - 1 ? n + NI * NJ : -1};
    const long fI0 = (k * NJ + j) * (NI + 1) + i, fJ0 = (k * (NJ + 1) + j) * NI + i, fK0 = (k * NJ + j) * NI + i;
    double g[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        double sx, sy, sz;
        if (q == 0) { sx = -c.six[fI0]; sy = -c.siy[fI0]; sz = -c.siz[fI0]; }
        else if (q == 1) { sx = c.six[fI0 + 1]; sy = c.siy[fI0 + 1]; sz = c.siz[fI0 + 1]; }
        else if (q == 2) { sx = c.sjx[fJ0 + NI]; sy = c.sjy[fJ0 + NI]; sz = c.sjz[fJ0 + NI]; }
        else if (q == 3) { sx = -c.sjx[fJ0]; sy = -c.sjy[fJ0]; sz = -c.sjz[fJ0]; }
        else if (q == 4) { sx = -c.skx[fK0]; sy = -c.sky[fK0]; sz = -c.skz[fK0]; }
        else { sx = c.skx[fK0 + NI * NJ]; sy = c.sky[fK0 + NI * NJ]; sz = c.skz[fK0 + NI * NJ]; }
        double fv[4] = {fc[0], fc[1], fc[2], fc[3]};
        if (nb[q] >= 0) {
            double fn[4];
            visc_fields(c, U, nb[q], fn);
            const double wq = c.gw[q * N + n];
#pragma unroll
            for (int v = 0; v < 4; ++v) fv[v] = fc[v] + (fn[v] - fc[v]) * wq;
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            g[v][0] = g[v][0] + sx * fv[v];
            g[v][1] = g[v][1] + sy * fv[v];
            g[v][2] = g[v][2] + sz * fv[v];
        }
    }
    const double iv = 1.0 / c.vol[n];
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int d = 0; d < 3; ++d) c.grad[(3 * v + d) * N + n] = g[v][d] * iv;
}

// Solver::computeViscousFlux (solver.cpp:2173-2226) without eddy viscosity
// (turbulence_model_none.hpp:49-51) and computeViscousStressTensor (math_utils.cpp:593-608),
// per unit area through S/|S|
__device__ inline void viscous_flux(const CaseDev& c, const double Ua[5], const double g[12], double Sx, double Sy,
                                    double Sz, double F[5]) {
    double W[5];
    W[0] = Ua[0];
#pragma unroll
    for (int v = 1; v < 5; ++v) W[v] = Ua[v] / Ua[0];
    const Gas<0> gs(c);
    const double T = pressure_w<double, 0>(gs, W) / (W[0] * c.rgas);
    const double muL = visc_mu(c, T);          // one law evaluation for mu and kappa
    const double mu = muL + 0.0;                // + eddy viscosity (none)
    const double kap = c.cp_v * muL / c.pr + 0.0;
    const double lam = -2.0 / 3.0 * mu;
    const double div = g[0] + g[4] + g[8];
    const double txx = lam * div + 2.0 * mu * g[0];
    const double tyy = lam * div + 2.0 * mu * g[4];
    const double tzz = lam * div + 2.0 * mu * g[8];
    const double txy = mu * (g[1] + g[3]), txz = mu * (g[2] + g[6]), tyz = mu * (g[5] + g[7]);
    const double thx = txx * W[1] + txy * W[2] + txz * W[3] + kap * g[9];
    const double thy = txy * W[1] + tyy * W[2] + tyz * W[3] + kap * g[10];
    const double thz = txz * W[1] + tyz * W[2] + tzz * W[3] + kap * g[11];
    const double m = sqrt(Sx * Sx + Sy * Sy + Sz * Sz);
    const double nx = Sx / m, ny = Sy / m, nz = Sz / m;
    F[0] = 0.0;
    F[1] = txx * nx + txy * ny + txz * nz;
    F[2] = txy * nx + tyy * ny + tyz * nz;
    F[3] = txz * nx + tyz * ny + tzz * nz;
    F[4] = thx * nx + thy * ny + thz * nz;
}

// Solver::computeViscousFluxResiduals (solver.cpp:2060-2170): the internal faces before
// each node (I, J, and K in 3-D), state and gradients averaged over the two nodes; stores
// flux * |S|.  Work: one entry per 256 nodes of a viscous case.
__global__ void __launch_bounds__(256) k_vface(Launch La) {
    const Work w = La.work[blockIdx.x];
    if (w.cs < 0) return;   // padding entry of an XCD-ordered table
    const CaseDev& c = La.cases[w.cs];
    if (*c.stop) return;
    const long n = (long)w.x * 256 + threadIdx.x, N = c.N;
    if (n >= N) return;
    const long NI = c.ni, NJ = c.nj;
    const long i = n % NI, j = (n / NI) % NJ, k = n / (NI * NJ);
    const double* __restrict__ U = stage_input(c, La);
    const int nd = (c.topo == TOPO_3D) ? 3 : 2;
    // the node's own state and gradients, loaded once for its two or three faces
    double Un[5], Gn[12];
#pragma unroll
    for (int v = 0; v < 5; ++v) Un[v] = U[v * N + n];
#pragma unroll
    for (int q = 0; q < 12; ++q) Gn[q] = c.grad[q * N + n];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const long idx = (d == 0) ? i : (d == 1 ? j : k);
        if (d >= nd || idx == 0) continue;
        const long l = n - ((d == 0) ? 1 : (d == 1 ? NI : NI * NJ));
        double Sx, Sy, Sz;
        if (d == 0) { const long f = (k * NJ + j) * (NI + 1) + i; Sx = c.six[f]; Sy = c.siy[f]; Sz = c.siz[f]; }
        else if (d == 1) { const long f = (k * (NJ + 1) + j) * NI + i; Sx = c.sjx[f]; Sy = c.sjy[f]; Sz = c.sjz[f]; }
        else { const long f = (k * NJ + j) * NI + i; Sx = c.skx[f]; Sy = c.sky[f]; Sz = c.skz[f]; }
        double Ua[5], g[12], F[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) Ua[v] = (U[v * N + l] + Un[v]) * 0.5;
#pragma unroll
        for (int q = 0; q < 12; ++q) g[q] = (c.grad[q * N + l] + Gn[q]) * 0.5;
        viscous_flux(c, Ua, g, Sx, Sy, Sz, F);
        const double m = sqrt(Sx * Sx + Sy * Sy + Sz * Sz);
#pragma unroll
        for (int e = 0; e < 4; ++e) c.vf[(4 * d + e) * N + n] = F[e + 1] * m;
    }
}

// The viscous faces of node (i, j, k) added to its advection residual: as the right node
// of face i it receives +flux*|S|, as the left node of face i+1 -flux*|S|
// (residuals.subtract / add of flux*|S|*(-1)), I then J then K.
__device__ __forceinline__ void viscous_add(const CaseDev& c, int i, int j, int k, long n, double R[5]) {
    const long N = c.N, NI = c.ni, NJ = c.nj;
    const double* __restrict__ f = c.vf;
    const bool k3 = (c.topo == TOPO_3D);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        double r = R[e + 1];
        if (i > 0) r = r + f[e * N + n];
        if (i + 1 < NI) r = r - f[e * N + n + 1];
        if (j > 0) r = r + f[(4 + e) * N + n];
        if (j + 1 < NJ) r = r - f[(4 + e) * N + n + NI];
        if (k3) {
            if (k > 0) r = r + f[(8 + e) * N + n];
            if (k + 1 < c.nk) r = r - f[(8 + e) * N + n + NI * NJ];
        }
        R[e + 1] = r;
    }
}

}  // namespace dev
}  // namespace tbfm
