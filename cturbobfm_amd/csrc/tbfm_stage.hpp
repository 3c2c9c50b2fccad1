// Kernels of one explicit iteration: time step, radial-equilibrium profiles, boundary
// fluxes, the fused residual + RK stage kernel, periodic closure and norm reductions.
#pragma once
#include "tbfm_physics.hpp"
#include "tbfm_viscous.hpp"

namespace tbfm {
namespace dev {

// Node and state offsets are 32-bit: the library rejects grids with 5*N >= 2^31.
template <typename CT>
__device__ __forceinline__ int nid(const CT& c, int i, int j, int k) {
    return (k * c.nj + j) * c.ni + i;
}

__device__ __forceinline__ void load5(const double* U, int N, int n, double out[5]) {
#pragma unroll
    for (int v = 0; v < 5; ++v) out[v] = gld(U + (long)v * N, (unsigned)n);
}

// rotateStateVectorAlongXAxis (src/math_utils.cpp:109-122)
__device__ __forceinline__ void rotx(double s[5], double cs, double sn) {
    double y = s[2] * cs - s[3] * sn;
    double z = s[2] * sn + s[3] * cs;
    s[2] = y;
    s[3] = z;
}

// Node gk of the (i,j) line along k, including the two virtual planes used by the face
// loop: -1 and NK (periodic wrap with rotation, solver.cpp:1964-1979, or linear
// extrapolation, solver.cpp:1980-1993).
// The source planes of a virtual plane are written relative to gk (not as the constants
// 0, 1, NK-2, NK-1) so that the compiler does not hoist their addresses out of the k loop
// into registers that the fused kernel needs.
__device__ inline void load_k(const CaseDev& c, const double* U, int i, int j, int gk, double out[5]) {
    const int NK = c.nk;
    const int N = (int)c.N;
    const int col = j * c.ni + i, NIJ = c.ni * c.nj;
    if (gk >= 0 && gk < NK) {
        load5(U, N, gk * NIJ + col, out);
    } else if (c.periodic) {
        const bool lo = gk < 0;
        load5(U, N, (lo ? gk + NK - 1 : gk - NK + 1) * NIJ + col, out);
        rotx(out, lo ? c.per_cm : c.per_c, lo ? c.per_sm : c.per_s);
    } else {
        // gk = -1: a = plane 0, b = plane 1; gk = NK: a = plane NK-1, b = plane NK-2
        const int d = (gk < 0) ? 1 : -1;
        double a[5], b[5];
        load5(U, N, (gk + d) * NIJ + col, a);
        load5(U, N, (gk + 2 * d) * NIJ + col, b);
#pragma unroll
        for (int v = 0; v < 5; ++v) out[v] = (gk < 0) ? a[v] - (b[v] - a[v]) : a[v] + (a[v] - b[v]);
    }
}

// ------------------------------------------------------------------ node sources
// Geometric (3-D blade rows: solver.cpp:2855-2867; axisymmetric: :2833-2852), Gong
// (:2870-2876, :3074-3145) and body-force (:2878-2893) sources of one node.  The stage
// kernel subtracts the three vectors from R in that order; an absent term is +0, and
// x - (+0) == x for every x, so the order and the values are the reference's.
// Also writes the body-force side outputs (inviscid / viscous force, deviation angle)
// on the last stage.
// Rotational speed of slab node s: rpm 2 pi / 60 times ROTATIONAL_SPEED_SCALING_FACTOR, or
// with ACCELERATION_ACTIVE times the factor of the iteration's physical time: the linear
// ramp between ACCELERATION_INITIAL_TIME and _FINAL_TIME, the constant factor outside
// (Config::getRotationalSpeedScalingFactor(t), config.cpp:695-718).
__device__ __forceinline__ double slab_omega(const CaseDev& c, long s) {
    if (!c.accel) return c.s_om[s];
    const double t = *c.t_now;
    double f;
    if (t < c.acc_t0 || t > c.acc_t1) f = c.om_scaling;
    else f = c.acc_s0 + (c.acc_s1 - c.acc_s0) * (t - c.acc_t0) / (c.acc_t1 - c.acc_t0);
    double om = c.s_om[s];
    om *= f;
    return om;
}

// SourceBFMBase::computePerturbationFactor (source_bfm_base.cpp:215-243) of slab node s at
// the iteration's physical time: the spatial Gaussian (host) times the temporal one about
// the middle of the window (sigma a third of its duration), times the scaling factor.
__device__ __forceinline__ double perturbation_factor(const CaseDev& c, long s) {
    if (!c.pert) return 0.0;
    const double t = *c.t_now;
    if (t < c.pert_start || t > c.pert_start + c.pert_dur) return 0.0;
    const double space = c.s_psp[s];
    if (space < 0.0) return 0.0;
    const double peak = c.pert_start + c.pert_dur / 2.0;
    const double sigma = c.pert_dur / 3.0;
    const double time = exp(-(pow(t - peak, 2.0)) / pow(sigma, 2.0));
    return c.pert_scale * space * time;
}

// Geometric source of one node (axisymmetric: (rho u_theta^2 + p, -rho u_r u_theta) A_K
// dtheta; 3-D blade rows: the same over r times V), in the source type T.  col = j*NI + i.
template <int FL, typename T>
__device__ inline void geometric_source(const CaseDev& c, const Gas<FL>& gs, const double* Wd, long s, long col,
                                        double Vd, double g[5]) {
    const T Wc[5] = {T(Wd[0]), T(Wd[1]), T(Wd[2]), T(Wd[3]), T(Wd[4])};
    const T V = T(Vd);
    T p = pressure_w<T, FL>(gs, Wc);
    const T cs = T(c.s_cs[s]), sn = T(c.s_sn[s]);
    V3T<T> vcy = cyl_of(v3(Wc[1], Wc[2], Wc[3]), cs, sn);
    V3T<T> scy;
    if (c.topo == TOPO_AXI) {
        const T gf = T(c.axi_gf[col]);
        scy = v3(T(0.0), (+Wc[0] * vcy.z * vcy.z + p) * gf, -Wc[0] * vcy.y * vcy.z * gf);
        V3T<T> sca = cart_of(scy, cs, sn);
        g[0] = 0.0; g[1] = sca.x; g[2] = sca.y; g[3] = sca.z; g[4] = 0.0;
    } else {
        const T r = T(c.s_r[s]);
        scy = v3(T(0.0), (+Wc[0] * vcy.z * vcy.z + p) / r, -Wc[0] * vcy.y * vcy.z / r);
        V3T<T> sca = cart_of(scy, cs, sn);
        g[0] = T(0.0) * V; g[1] = sca.x * V; g[2] = sca.y * V; g[3] = sca.z * V; g[4] = T(0.0) * V;
    }
}

template <bool THREED, int FL>
__device__ inline void source_terms(const CaseDev& c, const Gas<FL>& gs, const double* __restrict__ Uin, int last, int i, int j,
                                    int k, int n, long s, const double* Wc, double g[5], double gong[5],
                                    double bf[5]) {
    const int NI = c.ni, NJ = c.nj, NK = c.nk;
    const int N = (int)c.N;
#pragma unroll
    for (int v = 0; v < 5; ++v) { g[v] = 0.0; gong[v] = 0.0; bf[v] = 0.0; }
    const unsigned char fl = c.sflag[s];
    const double V = c.vol[n];
    const bool present = (fl & 1) != 0;
    const bool geo = (c.topo == TOPO_AXI) || (THREED && present);
    if (geo) {
        // mixed precision: in FP32 (ideal gas; the table fluid is FP64 only)
        if constexpr (FL == 0) {
            if (c.fp32) geometric_source<FL, float>(c, gs, Wc, s, (long)j * NI + i, V, g);
            else geometric_source<FL, double>(c, gs, Wc, s, (long)j * NI + i, V, g);
        } else {
            geometric_source<FL, double>(c, gs, Wc, s, (long)j * NI + i, V, g);
        }
    }
    const double om = slab_omega(c, s);
    if (THREED && present) {
        if (!(fabs(om) < 1E-3)) {
            double U0g[5], U1g[5], U2g[5];
            int k0g = k, k1, k2;
            if (om > 0) {
                k1 = (k == 0) ? NK - 2 : (k == 1 ? 0 : k - 1);
                k2 = (k == 0) ? NK - 3 : (k == 1 ? NK - 2 : k - 2);
            } else {
                k1 = k + 1; k2 = k + 2;
                if (k == NK - 1) { k0g = 0; k1 = 1; k2 = 2; }
                else if (k == NK - 2) { k1 = 0; k2 = 1; }
            }
            load5(Uin, N, nid(c, i, j, k0g), U0g);
            load5(Uin, N, nid(c, i, j, k1), U1g);
            load5(Uin, N, nid(c, i, j, k2), U2g);
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                double d;
                if (om > 0.0) d = (U0g[v] * 3.0 - U1g[v] * 4.0 + U2g[v]) / (2 * c.dtheta);
                else d = (U0g[v] * (-3.0) + U1g[v] * 4.0 - U2g[v]) / (2 * c.dtheta);
                gong[v] = (-d * om) * V;
            }
        }
    }
    if (c.bfm_active && (fl & 2)) {
        double fi[3], fv[3], dv;
        bool keepv;
        // mixed precision: the body force in FP32 (ideal gas; the table fluid is FP64 only)
        const double pf = perturbation_factor(c, s);
        if constexpr (FL == 0) {
            if (c.fp32) bfm_source<FL, float>(c, gs, s, om, pf, Wc, V, bf, fi, fv, dv, keepv);
            else bfm_source<FL, double>(c, gs, s, om, pf, Wc, V, bf, fi, fv, dv, keepv);
        } else {
            bfm_source<FL, double>(c, gs, s, om, pf, Wc, V, bf, fi, fv, dv, keepv);
        }
        if (last && (fl & 4)) {
            c.finv[3 * n] = fi[0]; c.finv[3 * n + 1] = fi[1]; c.finv[3 * n + 2] = fi[2];
            c.dev[n] = dv;
        }
        // a stalled Hall-Thollet node keeps the viscous force of its latest unstalled
        // evaluation, which may be an earlier stage's: with the stall model on, every
        // stage stores it
        if ((fl & 4) && !keepv && (last || c.stall)) {
            c.fvisc[3 * n] = fv[0]; c.fvisc[3 * n + 1] = fv[1]; c.fvisc[3 * n + 2] = fv[2];
        }
    }
}

// Source terms of every node of the source slab (i in [bi0, bi1)) from the stage input,
// stored as [geometric | Gong | body force][5][ns] for the stage kernel.  One thread per
// slab node, run before the stage kernel: the transcendental-heavy blade-row work runs
// at full occupancy and without diverging the stencil kernel's waves.  Work: one entry
// per 256 slab nodes of a case with sources.
template <bool THREED, int FL>
__device__ inline void sources_block(const CaseDev& c, const Launch& La, int bx) {
    const int s = bx * 256 + threadIdx.x;
    if (s >= c.ns) return;
    const int SW = c.bi1 - c.bi0, NJ = c.nj;
    const int i = c.bi0 + s % SW, j = (s / SW) % NJ, k = s / (SW * NJ);
    const int n = nid(c, i, j, k);
    const double* Uin = stage_input(c, La);
    const Gas<FL> gs(c);
    double U[5], W[5];
    load5(Uin, (int)c.N, n, U);
    W[0] = U[0];
#pragma unroll
    for (int v = 1; v < 5; ++v) W[v] = U[v] / U[0];
    double g[5], go[5], bf[5];
    source_terms<THREED, FL>(c, gs, Uin, La.last, i, j, k, n, s, W, g, go, bf);
    // the geometric source's density, x-momentum and energy components are +0 (its
    // cylindrical source has no axial part): only its y and z components are kept
    const int ns = c.ns;
    if (c.fp32) {
        // FP32 flux: the sources as floats (the geometric and body-force ones are FP32
        // values already, the Gong stencil is rounded)
        float* sb = reinterpret_cast<float*>(c.srcbuf);
        sb[s] = float(g[2]);
        sb[ns + s] = float(g[3]);
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            sb[(2 + v) * ns + s] = float(go[v]);
            sb[(7 + v) * ns + s] = float(bf[v]);
        }
        return;
    }
    c.srcbuf[s] = g[2];
    c.srcbuf[ns + s] = g[3];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        c.srcbuf[(2 + v) * ns + s] = go[v];
        c.srcbuf[(7 + v) * ns + s] = bf[v];
    }
}

// --------------------------------------------------------------------- time step
// Solver::computeTimestepArray, inviscid (src/solver.cpp:1531-1630)
// One block of 256 nodes (bx = block index within the case); writes the nodes' time
// steps to dt_out and, for the global time step (or the turbomachinery time), the
// block's minimum to part_out[bx].
template <int FL>
__device__ inline void timestep_block(const CaseDev& c, const double* U, int bx, double* dt_out,
                                      double* part_out) {
    __shared__ double smin[256];
    const long n = (long)bx * 256 + threadIdx.x;
    double dtv = 1.0e300;
    if (n < c.N) {
        if (c.dt_mode == 2) {
            dtv = c.fixed_dt;
        } else {
            const int i = (int)(n % c.ni);
            const int j = (int)((n / c.ni) % c.nj);
            const int k = (int)(n / ((long)c.ni * c.nj));
            const long DI = c.ni + 1, DJ = c.nj + 1;
            const long d0 = ((long)k * DJ + j) * DI + i;
            const long dI = d0 + 1, dJ = d0 + DI, dK = d0 + DI * DJ;
            const double *X = c.dux, *Y = c.duy, *Z = c.duz;
            const unsigned u0 = (unsigned)d0, uI = (unsigned)dI, uJ = (unsigned)dJ, uK = (unsigned)dK;
            double eIx = gldnt(X, uI) - gldnt(X, u0), eIy = gldnt(Y, uI) - gldnt(Y, u0), eIz = gldnt(Z, uI) - gldnt(Z, u0);
            double eJx = gldnt(X, uJ) - gldnt(X, u0), eJy = gldnt(Y, uJ) - gldnt(Y, u0), eJz = gldnt(Z, uJ) - gldnt(Z, u0);
            double eKx = gldnt(X, uK) - gldnt(X, u0), eKy = gldnt(Y, uK) - gldnt(Y, u0), eKz = gldnt(Z, uK) - gldnt(Z, u0);
            double dsI = sqrt(eIx * eIx + eIy * eIy + eIz * eIz);
            double dsJ = sqrt(eJx * eJx + eJy * eJy + eJz * eJz);
            double dsK = sqrt(eKx * eKx + eKy * eKy + eKz * eKz);
            const Gas<FL> gs(c);
            double Un[5], W[5], p, rf;
            load5(U, c.N, n, Un);
            derive<double, FL>(Un, W, p, rf, gs);
            const double a = sound_speed_w<FL>(gs, W, p);
            double ui = fabs(W[1] * (eIx / dsI) + W[2] * (eIy / dsI) + W[3] * (eIz / dsI));
            double lam = (ui + a) / dsI;
            if (c.topo != TOPO_1D) {
                double uj = fabs(W[1] * (eJx / dsJ) + W[2] * (eJy / dsJ) + W[3] * (eJz / dsJ));
                lam += (uj + a) / dsJ;
            }
            if (c.topo == TOPO_3D) {
                double uk = fabs(W[1] * (eKx / dsK) + W[2] * (eKy / dsK) + W[3] * (eKz / dsK));
                lam += (uk + a) / dsK;
            }
            double invH2 = 1.0 / (dsI * dsI);
            double mu = 0.0;
            if (c.visc) {
                if (c.topo != TOPO_1D) invH2 += 1.0 / (dsJ * dsJ);
                if (c.topo == TOPO_3D) invH2 += 1.0 / (dsK * dsK);
                mu = visc_mu(c, p / (W[0] * c.rgas));
            }
            double lv = 4.0 * (mu / W[0] + 0.0 / W[0]) * invH2;
            dtv = c.cfl / (lam + lv);
        }
        dt_out[n] = dtv;
    }
    if (c.dt_mode == 1 || c.turbo_on) {
        smin[threadIdx.x] = dtv;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (threadIdx.x < s) smin[threadIdx.x] = dmind(smin[threadIdx.x], smin[threadIdx.x + s]);
            __syncthreads();
        }
        if (threadIdx.x == 0) part_out[bx] = smin[0];
    }
}

// ------------------------------------------------------- turbomachinery history
// Solver::updateMassFlows + updateTurboPerformance (src/solver.cpp:1633-1781) of the
// state at the start of the iteration, and the physical time after it
// (solver.cpp:1383-1385: += the minimum time step).  Partial sums over 256 boundary
// faces run as blocks of k_turbo_partial; k_min_reduce adds them up (block tree reductions:
// the reference sums face by face, the two differ in the last bits only) and writes row
// *c.iter of turbo_hist.
constexpr int TURBO_SUMS = 5;   // per boundary: mass flow, rho u.S times p_t, T_t, h_t, s

// Block bx: boundary bx / turbo_nblk, faces [256 (bx % turbo_nblk), +256).
template <int FL>
__device__ inline void turbo_partial_block(const CaseDev& c, const double* U, int bx) {
    __shared__ double red[TURBO_SUMS][256];
    const int b = bx / c.turbo_nblk, f = (bx % c.turbo_nblk) * 256 + threadIdx.x;
    double acc[TURBO_SUMS] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (f < c.turbo_nf) {
        const Gas<FL> gs(c);
        const int n = c.turbo_node[b * c.turbo_nf + f];
        const double* S = c.turbo_S + 3 * ((long)b * c.turbo_nf + f);
        double Un[5];
        load5(U, (int)c.N, n, Un);
        const double rho = Un[0], ux = Un[1] / rho, uy = Un[2] / rho, uz = Un[3] / rho, et = Un[4] / rho;
        acc[0] = S[0] * Un[1] + S[1] * Un[2] + S[2] * Un[3];
        double pt, Tt, ht = 0.0, sv = 0.0;
        if constexpr (FL == 1) {
            const double qq = ux * ux + uy * uy + uz * uz;
            const double e = et - 0.5 * qq;
            pt = lut_total_pressure(*gs.L, rho, qq, et);
            Tt = lut_total_temperature(*gs.L, rho, qq, et);
            ht = et + lut_at(*gs.L, LUT_P, rho, e) / rho;
            sv = lut_at(*gs.L, LUT_S, rho, e);
        } else {
            const double gm = gs.g;
            const double um = sqrt(ux * ux + uy * uy + uz * uz);
            const double e = et - 0.5 * (um * um);
            const double pr = (gm - 1) * rho * e;
            const double a = sqrt(gm * pr / rho);
            const double M = um / a;
            pt = pr * pow(1.0 + (gm - 1.0) / 2.0 * M * M, gm / (gm - 1));
            const double T = pr / (rho * c.rgas);
            Tt = T * (1.0 + (gm - 1.0) / 2.0 * M * M);
        }
        acc[1] = S[0] * (rho * ux * pt) + S[1] * (rho * uy * pt) + S[2] * (rho * uz * pt);
        acc[2] = S[0] * (rho * ux * Tt) + S[1] * (rho * uy * Tt) + S[2] * (rho * uz * Tt);
        acc[3] = S[0] * (rho * ux * ht) + S[1] * (rho * uy * ht) + S[2] * (rho * uz * ht);
        acc[4] = S[0] * (rho * ux * sv) + S[1] * (rho * uy * sv) + S[2] * (rho * uz * sv);
    }
#pragma unroll
    for (int q = 0; q < TURBO_SUMS; ++q) red[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st)
#pragma unroll
            for (int q = 0; q < TURBO_SUMS; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < TURBO_SUMS) c.turbo_part[(long)bx * TURBO_SUMS + threadIdx.x] = red[threadIdx.x][0];
}

// Work: one entry per 256 turbo-boundary faces of a case with TURBO_OUTPUT (a launch of
// its own: the shared arrays would cost the time-step blocks occupancy).
__global__ void __launch_bounds__(256) k_turbo_partial(Launch L) {
    const Work w = L.work[blockIdx.x];
    const CaseDev& c = L.cases[w.cs];
    if (*c.stop) return;
    const double* U = c.U[L.in];
    if (c.real) turbo_partial_block<1>(c, U, w.x);
    else turbo_partial_block<0>(c, U, w.x);
}

// The boundary totals, the performance figures and the row; one block (256 threads,
// turbo_nblk <= 256 partials per boundary).
__device__ inline void turbo_final(const CaseDev& c) {
    __shared__ double red[2 * TURBO_SUMS][256];
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int q = 0; q < TURBO_SUMS; ++q)
            red[b * TURBO_SUMS + q][threadIdx.x] =
                (threadIdx.x < c.turbo_nblk) ? c.turbo_part[((long)b * c.turbo_nblk + threadIdx.x) * TURBO_SUMS + q] : 0.0;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int q = 0; q < 2 * TURBO_SUMS; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double mf0 = red[0][0], mf1 = red[TURBO_SUMS][0];
    double m = 0.5 * (fabs(mf0) + fabs(mf1));
    m *= c.turbo_mscale;
    double ptm[2], ttm[2], htm[2], sm[2];
    for (int b = 0; b < 2; ++b) {
        double mfl = red[b * TURBO_SUMS][0];
        if (fabs(mfl) < 1e-12) mfl = (mfl >= 0.0 ? 1e-12 : -1e-12);
        ptm[b] = red[b * TURBO_SUMS + 1][0] / mfl;
        ttm[b] = red[b * TURBO_SUMS + 2][0] / mfl;
        htm[b] = red[b * TURBO_SUMS + 3][0] / mfl;
        sm[b] = red[b * TURBO_SUMS + 4][0] / mfl;
    }
    const double PR = (fabs(ptm[0]) > 1e-12) ? ptm[1] / ptm[0] : 1.0;
    const double TR = (fabs(ttm[0]) > 1e-12) ? ttm[1] / ttm[0] : 1.0;
    double eta;
    if (c.real) {
        // enthalpy-based total-to-total efficiency with h(p_t,out, s_in)
        const double dh = htm[1] - htm[0], dhs = lut_h_p_s(*c.lut, ptm[1], sm[0]) - htm[0];
        if (fabs(dh) < 1e-12) eta = 1.0;
        else if (dh >= 0.0) eta = dhs / dh;
        else eta = (fabs(dhs) < 1e-12) ? 1.0 : dh / dhs;
    } else {
        const double gm = c.gamma;
        const double dT = TR - 1.0, dTs = pow(PR, (gm - 1.0) / gm) - 1.0;
        if (fabs(dT) < 1e-12) eta = 1.0;
        else if (dT >= 0.0) eta = dTs / dT;
        else eta = (fabs(dTs) < 1e-12) ? 1.0 : dT / dTs;
    }
    *c.t_now = *c.turbo_time;   // the time the body force of this iteration sees
    const double t = *c.turbo_time + *c.dtmin;
    *c.turbo_time = t;
    const long it = *c.iter;
    if (it >= 1 && it <= c.conv.hist_len) {
        double* row = c.turbo_hist + (it - 1) * 5;
        row[0] = t; row[1] = m; row[2] = PR; row[3] = TR; row[4] = eta;
    }
}

// Work: one entry per case using the global time step.
__global__ void __launch_bounds__(256) k_min_reduce(Launch L) {
    __shared__ double smin[256];
    const Work w = L.work[blockIdx.x];
    const CaseDev& c = L.cases[w.cs];
    if (*c.stop) return;
    double m = 1.0e300;
    for (int q = threadIdx.x; q < c.nblk_dt; q += 256) m = dmind(m, c.partial_min[q]);
    smin[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) smin[threadIdx.x] = dmind(smin[threadIdx.x], smin[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) c.dtmin[0] = smin[0];
    // the iteration's launch (L.last) also closes the turbomachinery row
    if (c.turbo_on && L.last) {
        __syncthreads();
        turbo_final(c);
    }
}

// ------------------------------------------------------- radial equilibrium profile
// Solver::updateRadialProfiles + integrateRadialEquilibrium (src/solver.cpp:2688-2787,
// src/math_utils.cpp:141-170).  The hub pressure per profile: radial equilibrium keeps
// it; a throttle sets p0 of the first inlet + k_t m^2 with the turbomachinery mass flow;
// a mass-flow outlet relaxes it towards the target (error clamped to the maximum
// change, at least 100 Pa).  The boundary mass flows come from the partial sums of
// k_turbo_partial (launched before), added as k_min_reduce adds them.  The terms
// rho*ut*ut/r of the outlet column are evaluated by all threads in parallel; the
// trapezoid recurrence P[m] = P[m-1] + 0.5*(d[m-1] + d[m])*dr is then run by one
// thread, in order, exactly as the reference accumulates it.  One block per case with
// profiles.
__device__ inline void radial_profile_block(const CaseDev& c, const double* U) {
    __shared__ double dval[257];
    __shared__ double mfs[2][256];
    const int tid = threadIdx.x;
    double mb[2] = {0.0, 0.0}, mturbo = 0.0;
    if (c.prof_need_mf) {
        for (int b = 0; b < 2; ++b)
            mfs[b][tid] = (tid < c.turbo_nblk) ? c.turbo_part[((long)b * c.turbo_nblk + tid) * TURBO_SUMS] : 0.0;
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if (tid < st) {
                mfs[0][tid] += mfs[0][tid + st];
                mfs[1][tid] += mfs[1][tid + st];
            }
            __syncthreads();
        }
        mb[0] = mfs[0][0];
        mb[1] = mfs[1][0];
        mturbo = 0.5 * (fabs(mb[0]) + fabs(mb[1]));
        mturbo *= c.turbo_mscale;
    }
    // GreitzerModel::computePlenumPressure (greitzer_model.cpp:36-56), one step per
    // iteration; thread 0 alone uses the hub
    double gz_hub = 0.0;
    if (c.greitzer && tid == 0) {
        double* g = c.gz;
        const double a = c.real ? lut_at(*c.lut, LUT_A_PT, g[0], g[2]) : sqrt(c.gamma * c.rgas * g[2]);
        const double np = g[0] + c.gz_dt * a * a / c.gz_vol * (mturbo - g[1]);
        g[0] = np;
        const double dp = np - c.gz_amb;
        g[1] = (dp > 0.0) ? sqrt(dp / c.gz_kt) : 0.0;
        gz_hub = np;
    }
    for (int q = 0; q < c.nprof; ++q) {
        const int n = c.prof_n[q];
        double* P = c.prof_p + (long)q * c.nj;
        double hub = P[0];
        if (c.greitzer) {
            hub = gz_hub;
        } else if (c.prof_kind[q] == 1) {
            hub = c.prof_par[q][1] + c.prof_par[q][0] * mturbo * mturbo;
        } else if (c.prof_kind[q] == 2) {
            double cur = fabs(mb[c.prof_mfb[q]]);
            cur *= c.turbo_mscale;
            const double target = c.prof_par[q][0], relax = c.prof_par[q][1], maxc = c.prof_par[q][2];
            const double err = (cur - target) / target;
            double d = relax * err;
            d = (d < -maxc) ? -maxc : ((maxc < d) ? maxc : d);
            double cp = P[0];
            if (cp <= 0.0) cp = c.prof_par[q][3];
            hub = cp * (1.0 + d);
            hub = (hub < 100.0) ? 100.0 : hub;
        }
        if (!c.prof_radial[q] || n <= 1) {
            // uniform profile at the hub value (a mass-flow outlet clamps the face's j
            // into the profile: every entry of the row holds the value)
            if (tid == 0)
                for (int m = 0; m < (c.prof_kind[q] == 2 ? c.nj : n); ++m) P[m] = hub;
            __syncthreads();
            continue;
        }
        const int i = c.prof_i[q];
        const double* r = c.prof_r + (long)q * c.nj;
        const double* cs = c.prof_cs + (long)q * c.nj;
        const double* sn = c.prof_sn + (long)q * c.nj;
        if (tid == 0) P[0] = hub;
        for (int base = 0; base < n; base += 256) {
            // dval[0] carries d[base-1] of the previous chunk
            const int m = base + tid;
            if (m < n) {
                const int j = c.prof_j0[q] + m;
                double Un[5], W[5];
                load5(U, (int)c.N, nid(c, i, j, 0), Un);
                W[0] = Un[0];
                W[1] = Un[1] / Un[0];
                W[2] = Un[2] / Un[0];
                W[3] = Un[3] / Un[0];
                const double ut = fabs(-W[2] * sn[m] + W[3] * cs[m]);
                const double rho = W[0];
                dval[tid + 1] = rho * ut * ut / r[m];
            }
            __syncthreads();
            if (tid == 0) {
                const int top = min(n, base + 256);
                for (int mm = (base == 0 ? 1 : base); mm < top; ++mm) {
                    const double dr = r[mm] - r[mm - 1];
                    P[mm] = P[mm - 1] + 0.5 * (dval[mm - base] + dval[mm - base + 1]) * dr;
                }
                dval[0] = dval[top - base];
            }
            __syncthreads();
        }
        if (c.prof_kind[q] == 2 && tid == 0)
            for (int m = n; m < c.nj; ++m) P[m] = P[n - 1];
    }
}

// Start of an iteration: radial-equilibrium profiles (z = 1, one block per case with
// profiles) and the time step (z = 0, one block per 256 nodes), both from the state in
// role `in`; independent of each other, so they share one launch.
__global__ void __launch_bounds__(256) k_iterstart(Launch L) {
    const Work w = L.work[blockIdx.x];
    if (w.cs < 0) return;   // padding entry of an XCD-ordered table
    const CaseDev& c = L.cases[w.cs];
    if (*c.stop) return;
    const double* U = c.U[L.in];
    if (w.z == 1) radial_profile_block(c, U);
    else if (c.real) timestep_block<1>(c, U, w.x, c.dt, c.partial_min);
    else timestep_block<0>(c, U, w.x, c.dt, c.partial_min);
}

// ------------------------------------------------------------- boundary fluxes
// All boundary faces of all sides, one thread per face.  Start faces use the
// outward vector -S and the node at the face; end faces +S and the node before it
// (solver.cpp:1911-1957).  Result already multiplied by |S|.  Work: one entry per 256
// boundary faces of a case.
template <int FL>
__device__ inline void boundary_block(const CaseDev& c, const double* U, int bx) {
    const long NI = c.ni, NJ = c.nj, NK = c.nk;
    long t = (long)bx * 256 + threadIdx.x;
    const long nI = NJ * NK, nJ = NI * NK, nK = NI * NJ;
    const bool doJ = (c.topo != TOPO_1D), doK = (c.topo == TOPO_3D);
    int dir, side;
    long q;
    if (t < 2 * nI) { dir = 0; side = (int)(t / nI); q = t % nI; }
    else if (doJ && t < 2 * nI + 2 * nJ) { t -= 2 * nI; dir = 1; side = (int)(t / nJ); q = t % nJ; }
    else if (doK && t < 2 * nI + 2 * nJ + 2 * nK) { t -= 2 * nI + 2 * nJ; dir = 2; side = (int)(t / nK); q = t % nK; }
    else return;
    int i, j, k, fi, fj, fk;
    long sidx, nsz;
    const double *sx, *sy, *sz;
    const int* pid;
    const double* dirv;
    double* out;
    if (dir == 0) {
        j = (int)(q % NJ); k = (int)(q / NJ);
        fi = side ? (int)NI : 0; fj = j; fk = k;
        i = side ? (int)NI - 1 : 0;
        sidx = ((long)k * NJ + j) * (NI + 1) + fi;
        sx = c.six; sy = c.siy; sz = c.siz; pid = c.pI; dirv = c.dirI; out = c.bnd; nsz = nI;
    } else if (dir == 1) {
        i = (int)(q % NI); k = (int)(q / NI);
        fi = i; fj = side ? (int)NJ : 0; fk = k;
        j = side ? (int)NJ - 1 : 0;
        sidx = ((long)k * (NJ + 1) + fj) * NI + i;
        sx = c.sjx; sy = c.sjy; sz = c.sjz; pid = c.pJ; dirv = c.dirJ; out = c.bnd + c.boffJ; nsz = nJ;
    } else {
        i = (int)(q % NI); j = (int)(q / NI);
        fi = i; fj = j; fk = side ? (int)NK : 0;
        k = side ? (int)NK - 1 : 0;
        sidx = ((long)fk * NJ + j) * NI + i;
        sx = c.skx; sy = c.sky; sz = c.skz; pid = c.pK; dirv = c.dirK; out = c.bnd + c.boffK; nsz = nK;
    }
    double Sx = sx[sidx], Sy = sy[sidx], Sz = sz[sidx];
    if (!side) { Sx = -Sx; Sy = -Sy; Sz = -Sz; }
    double Un[5];
    load5(U, c.N, nid(c, i, j, k), Un);
    const int p = pid[side * nsz + q];
    const PatchDev pa = c.patch[p];
    double F[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const long it = *c.iter;
    const Gas<FL> gs(c);
    if (pa.type == PT_INLET && !side && c.nsw && c.nsw[nid(c, fi, fj, fk)]) {
        // an inlet face on a no-slip wall node: the wall's stagnation pressure
        // (boundary_inlet.cpp:16-19; end faces lie outside the node range and never match)
        const long w = nid(c, fi, fj, fk);
        stagnation_wall_flux(c, Sx, Sy, Sz, pa.v0, pa.v1, c.wallv[w], c.wallv[c.N + w], c.wallv[2 * c.N + w], F);
    } else if (pa.type == PT_INLET) {
        const double* d = dirv + (long)side * 3 * nsz;
        inlet_flux<FL>(c, gs, pa.tab, Un, Sx, Sy, Sz, pa.v0, pa.v1, d[q], d[nsz + q], d[2 * nsz + q], F);
    } else if (pa.type == PT_OUTLET) {
        outlet_flux<FL>(c, gs, Un, Sx, Sy, Sz, pa.v0, it, F);
    } else if (pa.type == PT_RADIAL) {
        outlet_flux<FL>(c, gs, Un, Sx, Sy, Sz, c.prof_p[(long)pa.profile * NJ + fj], it, F);
    } else if (pa.type == PT_WALL) {
        wall_flux<FL>(gs, Un, Sx, Sy, Sz, F);
    } else if (pa.type == PT_TRANSPARENT) {
        transparent_flux<FL>(c, gs, Un, Sx, Sy, Sz, F);
    }
    const double sm = sqrt(Sx * Sx + Sy * Sy + Sz * Sz);
    double* o = out + ((long)side * nsz + q) * 6;
#pragma unroll
    for (int v = 0; v < 5; ++v) o[v] = F[v] * sm;
}

// Bisection table of an ideal-gas inlet patch (p0, T0): node q of the first INLET_LEVELS
// levels, reached from the root by the halvings its heap index encodes; T and V of its
// midpoint exactly as the inlet flux computes them (boundary_base.cpp:51-68).
__global__ void __launch_bounds__(256) k_inlet_table(double* tab, double pt, double Tt, double g, double Rg) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= INLET_NODES) return;
    const double cp = g * Rg / (g - 1.0);
    const double htIn = cp * Tt;
    const double ex = (g - 1.0) / g;
    double pLow = 0.05 * pt, pHigh = 0.999999 * pt;
    const int level = 31 - __clz(q + 1);
    const int pos = q + 1 - (1 << level);
    for (int s = level - 1; s >= 0; --s) {
        const double pMid = 0.5 * (pLow + pHigh);
        if ((pos >> s) & 1) pHigh = pMid; else pLow = pMid;
    }
    const double pMid = 0.5 * (pLow + pHigh);
    const double TMid = Tt * pow(pMid / pt, ex);
    const double hMid = cp * TMid;
    tab[q] = TMid;
    tab[INLET_NODES + q] = sqrt(2.0 * dmaxd(0.0, htIn - hMid));
}

// Ahead of each stage kernel: boundary-face fluxes (z = 0, one block per 256 boundary
// faces) and slab source terms (z = 1: 3-D case, z = 2: 2-D / axisymmetric; one block per
// 256 slab nodes), all from the stage input.  Independent work in one launch: the
// latency-bound boundary blocks (inlet bisection) overlap the throughput-bound sources.
// One instance per fluid (the table fluid's boundary states would otherwise cost the
// ideal-gas blocks occupancy); each skips the cases of the other fluid.
template <int FL>
__global__ void __launch_bounds__(256) k_prestage(Launch La) {
    const Work w = La.work[blockIdx.x];
    if (w.cs < 0) return;   // padding entry of an XCD-ordered table
    const CaseDev& c = La.cases[w.cs];
    if (*c.stop || c.real != FL) return;
    if (w.z == 0) boundary_block<FL>(c, stage_input(c, La), w.x);
    else if (w.z == 1) sources_block<true, FL>(c, La, w.x);
    else sources_block<false, FL>(c, La, w.x);
}

}  // namespace dev
}  // namespace tbfm
