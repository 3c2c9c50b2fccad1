// The fused residual + RK stage kernel.
//
// One block owns a TI x TJ tile of (i, j) columns and marches along k over `kc` planes.
// Per plane it stages the conserved state, primitives, pressure and spectral factor of
// the tile plus a 2-node halo in LDS (ghost nodes at the domain ends are built exactly
// as the reference extrapolates them), evaluates every I and J face of the tile once,
// and takes the K faces from a per-thread register queue of the column (planes k-1..k+2).
// The residual of each node is then assembled in the reference's order
//   (I left, I right, J left, J right, K left, K right, geometric, Gong, BFM)
// (solver.cpp:1794-1821, 1997-2006, 2853-2891) and the RK update
// U = U0 - R * (alpha dt / V) (solver.cpp:2230-2249) is applied in place.  Nodes on the
// periodic planes k = 0 and k = NK-1 park their residual for the periodic-closure kernel.
#pragma once
#include "tbfm_stage.hpp"

namespace tbfm {
namespace dev {

template <int SCH>
struct PlaneLds {
    double U[5][NP];
    double W[5][NP];
    double P[NP];
    double RF[NP];
    double SI[TJ][TI + 2];      // JST I-sensor of nodes i0-1 .. i0+TI
    double SJ[TJ + 2][TI];      // JST J-sensor of rows j0-1 .. j0+TJ
    double FI[5][TJ][TI + 1];   // I-face flux*|S|: face i0+m of row j0+tj
    double FJ[5][TJ + 1][TI];   // J-face flux*|S|
    double red[NT];
};

// Fills LDS position `pos` of plane k: real node, ghost (one step outside the domain in
// i or j) or unused.
template <int SCH>
__device__ inline void fill_position(const CaseDev& c, const double* U, PlaneLds<SCH>& L, int pos, int i0, int j0, int k) {
    const int hi = pos % HI, hj = pos / HI;
    const int gi = i0 - 2 + hi, gj = j0 - 2 + hj;
    const int NI = c.ni, NJ = c.nj;
    const bool iin = (gi >= 0 && gi < NI), jin = (gj >= 0 && gj < NJ);
    double Un[5];
    bool ok = true;
    if (iin && jin) {
        load5(U, c.N, nid(c, gi, gj, k), Un);
    } else if (jin && (gi == -1 || gi == NI)) {
        double a[5], b[5];
        if (gi < 0) {
            load5(U, c.N, nid(c, 0, gj, k), a);
            load5(U, c.N, nid(c, 1, gj, k), b);
#pragma unroll
            for (int v = 0; v < 5; ++v) Un[v] = a[v] - (b[v] - a[v]);
        } else {
            load5(U, c.N, nid(c, NI - 1, gj, k), a);
            load5(U, c.N, nid(c, NI - 2, gj, k), b);
#pragma unroll
            for (int v = 0; v < 5; ++v) Un[v] = a[v] + (a[v] - b[v]);
        }
    } else if (iin && (gj == -1 || gj == NJ) && c.topo != TOPO_1D) {
        double a[5], b[5];
        if (gj < 0) {
            load5(U, c.N, nid(c, gi, 0, k), a);
            load5(U, c.N, nid(c, gi, 1, k), b);
#pragma unroll
            for (int v = 0; v < 5; ++v) Un[v] = a[v] - (b[v] - a[v]);
        } else {
            load5(U, c.N, nid(c, gi, NJ - 1, k), a);
            load5(U, c.N, nid(c, gi, NJ - 2, k), b);
#pragma unroll
            for (int v = 0; v < 5; ++v) Un[v] = a[v] + (a[v] - b[v]);
        }
    } else {
        ok = false;
    }
    if (!ok) {
#pragma unroll
        for (int v = 0; v < 5; ++v) { L.U[v][pos] = 1.0; L.W[v][pos] = 1.0; }
        L.P[pos] = 1.0;
        L.RF[pos] = 0.0;
        return;
    }
    double W[5], p, rf;
    derive<double>(Un, W, p, rf, c.gamma, c.gm1);
#pragma unroll
    for (int v = 0; v < 5; ++v) { L.U[v][pos] = Un[v]; L.W[v][pos] = W[v]; }
    L.P[pos] = p;
    L.RF[pos] = rf;
}

// Face flux*|S| between LDS positions l and r (with ll, rr) along i or j.
template <int SCH, typename FT>
__device__ __forceinline__ void lds_face(const CaseDev& c, const PlaneLds<SCH>& L, int ll, int l, int r, int rr,
                                         double sl, double sr, double sx, double sy, double sz, double out[5]) {
    if (SCH == 0) {
        double Ull[5], Ul[5], Ur[5], Urr[5], Wl[5], Wr[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            Ull[v] = L.U[v][ll]; Ul[v] = L.U[v][l]; Ur[v] = L.U[v][r]; Urr[v] = L.U[v][rr];
            Wl[v] = L.W[v][l]; Wr[v] = L.W[v][r];
        }
        jst_face<FT>(Ull, Ul, Ur, Urr, Wl, Wr, L.RF[l], L.RF[r], sl, sr, sx, sy, sz, c.gm1, out);
    } else {
        double Wll[5], Wl[5], Wr[5], Wrr[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            Wll[v] = L.W[v][ll]; Wl[v] = L.W[v][l]; Wr[v] = L.W[v][r]; Wrr[v] = L.W[v][rr];
        }
        roe_face<FT>(Wll, Wl, Wr, Wrr, c.muscl, c.limiter, sx, sy, sz, c.gamma, out);
    }
}

template <int SCH, bool THREED, typename FT>
__global__ void __launch_bounds__(NT) k_stage(CaseDev c, StageDev st) {
    __shared__ PlaneLds<SCH> L;
    if (*c.stop) return;
    const int tid = threadIdx.x;
    const int ti = tid % TI, tj = tid / TI;
    const int i0 = blockIdx.x * TI, j0 = blockIdx.y * TJ;
    const int i = i0 + ti, j = j0 + tj;
    const int NI = c.ni, NJ = c.nj, NK = c.nk;
    const bool act = (i < NI) && (j < NJ);
    const int kc0 = blockIdx.z * st.kc;
    const int kc1 = min(NK, kc0 + st.kc);
    const bool doJ = (c.topo != TOPO_1D);
    const long N = c.N;
    const int own = (tj + 2) * HI + (ti + 2);
    const bool bladecol = THREED && c.bfm_active && act && c.blade2d[(long)j * NI + i];
    const long it = *c.iter;

    // K queue: slots 0..3 hold planes k-1, k, k+1, k+2 at the top of step k
    double qU[4][5], qW[4][5], qP[4], qRF[4];
    double FKl[5] = {0, 0, 0, 0, 0};
    if (THREED && act) {
        for (int s = 0; s < 4; ++s) {
            const int gk = kc0 - 2 + s;
            if (gk >= -1 && gk <= NK) {
                load_k(c, st.Uin, i, j, gk, qU[s]);
                derive<double>(qU[s], qW[s], qP[s], qRF[s], c.gamma, c.gm1);
            } else {
#pragma unroll
                for (int v = 0; v < 5; ++v) { qU[s][v] = 1.0; qW[s][v] = 1.0; }
                qP[s] = 1.0; qRF[s] = 0.0;
            }
        }
        // K face on the left of node kc0
        if (kc0 == 0) {
            if (!bladecol) {
                const long q = (long)j * NI + i;
#pragma unroll
                for (int v = 0; v < 5; ++v) FKl[v] = c.bK[(long)v * NI * NJ + q];
            }
        } else if (!bladecol) {
            const long sidx = ((long)kc0 * NJ + j) * NI + i;
            if (SCH == 0) {
                double sl = jst_sensor<double>(qP[0], qP[1], qP[2]);
                double sr = jst_sensor<double>(qP[1], qP[2], qP[3]);
                jst_face<FT>(qU[0], qU[1], qU[2], qU[3], qW[1], qW[2], qRF[1], qRF[2], sl, sr,
                             c.skx[sidx], c.sky[sidx], c.skz[sidx], c.gm1, FKl);
            } else {
                roe_face<FT>(qW[0], qW[1], qW[2], qW[3], c.muscl, c.limiter, c.skx[sidx], c.sky[sidx],
                             c.skz[sidx], c.gamma, FKl);
            }
        }
    }
    double nrm[5] = {0, 0, 0, 0, 0};

    for (int k = kc0; k < kc1; ++k) {
        double Uc[5], Wc[5], Pc = 0.0, RFc = 0.0;
        if (THREED) {
            if (act) {
                // shift the queue and bring in plane k+2
#pragma unroll
                for (int s = 0; s < 3; ++s) {
#pragma unroll
                    for (int v = 0; v < 5; ++v) { qU[s][v] = qU[s + 1][v]; qW[s][v] = qW[s + 1][v]; }
                    qP[s] = qP[s + 1];
                    qRF[s] = qRF[s + 1];
                }
                if (k + 2 <= NK) {
                    load_k(c, st.Uin, i, j, k + 2, qU[3]);
                    derive<double>(qU[3], qW[3], qP[3], qRF[3], c.gamma, c.gm1);
                }
#pragma unroll
                for (int v = 0; v < 5; ++v) { Uc[v] = qU[1][v]; Wc[v] = qW[1][v]; }
                Pc = qP[1];
                RFc = qRF[1];
            }
        } else if (act) {
            load5(st.Uin, N, nid(c, i, j, k), Uc);
            derive<double>(Uc, Wc, Pc, RFc, c.gamma, c.gm1);
        }
        // ---- stage plane k (own node from registers, halo by all threads)
        if (act) {
#pragma unroll
            for (int v = 0; v < 5; ++v) { L.U[v][own] = Uc[v]; L.W[v][own] = Wc[v]; }
            L.P[own] = Pc;
            L.RF[own] = RFc;
        }
        for (int pos = tid; pos < NP; pos += NT) {
            const int hi = pos % HI, hj = pos / HI;
            const bool inTile = (hi >= 2 && hi < TI + 2 && hj >= 2 && hj < TJ + 2);
            if (inTile && (i0 + hi - 2) < NI && (j0 + hj - 2) < NJ) continue;  // owner wrote it
            fill_position<SCH>(c, st.Uin, L, pos, i0, j0, k);
        }
        __syncthreads();
        // ---- JST sensors of the nodes the tile's faces touch
        if (SCH == 0) {
            for (int q = tid; q < TJ * (TI + 2); q += NT) {
                const int m = q % (TI + 2), r = q / (TI + 2);
                const int gi = i0 - 1 + m;
                double s = 0.0;
                if (gi >= 0 && gi < NI) {
                    const int pos = (r + 2) * HI + (m + 1);
                    s = jst_sensor<double>(L.P[pos - 1], L.P[pos], L.P[pos + 1]);
                }
                L.SI[r][m] = s;
            }
            if (doJ) {
                for (int q = tid; q < (TJ + 2) * TI; q += NT) {
                    const int m = q % TI, r = q / TI;
                    const int gj = j0 - 1 + r;
                    double s = 0.0;
                    if (gj >= 0 && gj < NJ) {
                        const int pos = (r + 1) * HI + (m + 2);
                        s = jst_sensor<double>(L.P[pos - HI], L.P[pos], L.P[pos + HI]);
                    }
                    L.SJ[r][m] = s;
                }
            }
            __syncthreads();
        }
        // ---- faces
        double FKr[5] = {0, 0, 0, 0, 0};
        if (act) {
            // I face on the left of the node (face index i)
            double F[5];
            if (i == 0) {
                const long q = (long)k * NJ + j;
#pragma unroll
                for (int v = 0; v < 5; ++v) F[v] = c.bI[(long)v * NJ * NK + q];
            } else {
                const long sidx = ((long)k * NJ + j) * (NI + 1) + i;
                lds_face<SCH, FT>(c, L, own - 2, own - 1, own, own + 1, (SCH == 0) ? L.SI[tj][ti] : 0.0,
                                  (SCH == 0) ? L.SI[tj][ti + 1] : 0.0, c.six[sidx], c.siy[sidx], c.siz[sidx], F);
            }
#pragma unroll
            for (int v = 0; v < 5; ++v) L.FI[v][tj][ti] = F[v];
            if (ti == TI - 1 || i == NI - 1) {
                if (i + 1 == NI) {
                    const long q = (long)k * NJ + j;
#pragma unroll
                    for (int v = 0; v < 5; ++v) F[v] = c.bI[(long)(5 + v) * NJ * NK + q];
                } else {
                    const long sidx = ((long)k * NJ + j) * (NI + 1) + i + 1;
                    lds_face<SCH, FT>(c, L, own - 1, own, own + 1, own + 2, (SCH == 0) ? L.SI[tj][ti + 1] : 0.0,
                                      (SCH == 0) ? L.SI[tj][ti + 2] : 0.0, c.six[sidx], c.siy[sidx], c.siz[sidx], F);
                }
#pragma unroll
                for (int v = 0; v < 5; ++v) L.FI[v][tj][ti + 1] = F[v];
            }
            if (doJ) {
                if (j == 0) {
                    const long q = (long)k * NI + i;
#pragma unroll
                    for (int v = 0; v < 5; ++v) F[v] = c.bJ[(long)v * NI * NK + q];
                } else {
                    const long sidx = ((long)k * (NJ + 1) + j) * NI + i;
                    lds_face<SCH, FT>(c, L, own - 2 * HI, own - HI, own, own + HI, (SCH == 0) ? L.SJ[tj][ti] : 0.0,
                                      (SCH == 0) ? L.SJ[tj + 1][ti] : 0.0, c.sjx[sidx], c.sjy[sidx], c.sjz[sidx], F);
                }
#pragma unroll
                for (int v = 0; v < 5; ++v) L.FJ[v][tj][ti] = F[v];
                if (tj == TJ - 1 || j == NJ - 1) {
                    if (j + 1 == NJ) {
                        const long q = (long)k * NI + i;
#pragma unroll
                        for (int v = 0; v < 5; ++v) F[v] = c.bJ[(long)(5 + v) * NI * NK + q];
                    } else {
                        const long sidx = ((long)k * (NJ + 1) + j + 1) * NI + i;
                        lds_face<SCH, FT>(c, L, own - HI, own, own + HI, own + 2 * HI, (SCH == 0) ? L.SJ[tj + 1][ti] : 0.0,
                                          (SCH == 0) ? L.SJ[tj + 2][ti] : 0.0, c.sjx[sidx], c.sjy[sidx], c.sjz[sidx], F);
                    }
#pragma unroll
                    for (int v = 0; v < 5; ++v) L.FJ[v][tj + 1][ti] = F[v];
                }
            }
            // K face on the right of the node (face index k+1)
            if (THREED && !bladecol) {
                if (k + 1 == NK) {
                    const long q = (long)j * NI + i;
#pragma unroll
                    for (int v = 0; v < 5; ++v) FKr[v] = c.bK[(long)(5 + v) * NI * NJ + q];
                } else {
                    const long sidx = ((long)(k + 1) * NJ + j) * NI + i;
                    if (SCH == 0) {
                        double sl = jst_sensor<double>(qP[0], qP[1], qP[2]);
                        double sr = jst_sensor<double>(qP[1], qP[2], qP[3]);
                        jst_face<FT>(qU[0], qU[1], qU[2], qU[3], qW[1], qW[2], qRF[1], qRF[2], sl, sr,
                                     c.skx[sidx], c.sky[sidx], c.skz[sidx], c.gm1, FKr);
                    } else {
                        roe_face<FT>(qW[0], qW[1], qW[2], qW[3], c.muscl, c.limiter, c.skx[sidx], c.sky[sidx],
                                     c.skz[sidx], c.gamma, FKr);
                    }
                }
            }
        }
        __syncthreads();
        // ---- residual assembly, sources, RK update
        if (act) {
            const long n = nid(c, i, j, k);
            double R[5];
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                double r = 0.0;
                r = (i == 0) ? r + L.FI[v][tj][ti] : r - L.FI[v][tj][ti];
                r = r + L.FI[v][tj][ti + 1];
                if (doJ) {
                    r = (j == 0) ? r + L.FJ[v][tj][ti] : r - L.FJ[v][tj][ti];
                    r = r + L.FJ[v][tj + 1][ti];
                }
                if (THREED && !bladecol) {
                    r = (k == 0) ? r + FKl[v] : r - FKl[v];
                    r = r + FKr[v];
                }
                R[v] = r;
            }
            // sources (solver.cpp:2790-2897)
            const bool srcOn = (c.topo == TOPO_AXI) || (THREED && c.bfm_active) || (c.topo == TOPO_2D && c.bfm_active);
            if (srcOn && i >= c.bi0 && i < c.bi1) {
                const long s = ((long)k * NJ + j) * (c.bi1 - c.bi0) + (i - c.bi0);
                const unsigned char fl = c.sflag[s];
                const double V = c.vol[n];
                const bool present = (fl & 1) != 0;
                const bool geo = (c.topo == TOPO_AXI) || (THREED && present);
                if (geo) {
                    double e = static_energy<double>(Wc[1], Wc[2], Wc[3], Wc[4]);
                    double p = c.gm1 * Wc[0] * e;
                    const double cs = c.s_cs[s], sn = c.s_sn[s];
                    V3 vcy = cyl_of(v3(Wc[1], Wc[2], Wc[3]), cs, sn);
                    V3 scy;
                    double g[5];
                    if (c.topo == TOPO_AXI) {
                        const double gf = c.axi_gf[(long)j * NI + i];
                        scy = v3(0.0, (+Wc[0] * vcy.z * vcy.z + p) * gf, -Wc[0] * vcy.y * vcy.z * gf);
                        V3 sca = cart_of(scy, cs, sn);
                        g[0] = 0.0; g[1] = sca.x; g[2] = sca.y; g[3] = sca.z; g[4] = 0.0;
                    } else {
                        const double r = c.s_r[s];
                        scy = v3(0.0, (+Wc[0] * vcy.z * vcy.z + p) / r, -Wc[0] * vcy.y * vcy.z / r);
                        V3 sca = cart_of(scy, cs, sn);
                        g[0] = 0.0 * V; g[1] = sca.x * V; g[2] = sca.y * V; g[3] = sca.z * V; g[4] = 0.0 * V;
                    }
#pragma unroll
                    for (int v = 0; v < 5; ++v) R[v] -= g[v];
                }
                if (THREED && present) {
                    // Gong source -Omega dU/dtheta, one-sided 2nd order (solver.cpp:3074-3145)
                    const double om = c.s_om[s];
                    if (!(fabs(om) < 1E-3)) {
                        double U0g[5], U1g[5], U2g[5];
                        if (om > 0) {
                            const int k1 = (k == 0) ? NK - 2 : (k == 1 ? 0 : k - 1);
                            const int k2 = (k == 0) ? NK - 3 : (k == 1 ? NK - 2 : k - 2);
                            load5(st.Uin, N, n, U0g);
                            load5(st.Uin, N, nid(c, i, j, k1), U1g);
                            load5(st.Uin, N, nid(c, i, j, k2), U2g);
                        } else {
                            int k0g = k, k1 = k + 1, k2 = k + 2;
                            if (k == NK - 1) { k0g = 0; k1 = 1; k2 = 2; }
                            else if (k == NK - 2) { k1 = 0; k2 = 1; }
                            load5(st.Uin, N, nid(c, i, j, k0g), U0g);
                            load5(st.Uin, N, nid(c, i, j, k1), U1g);
                            load5(st.Uin, N, nid(c, i, j, k2), U2g);
                        }
#pragma unroll
                        for (int v = 0; v < 5; ++v) {
                            double d;
                            if (om > 0.0) d = (U0g[v] * 3.0 - U1g[v] * 4.0 + U2g[v]) / (2 * c.dtheta);
                            else d = (U0g[v] * (-3.0) + U1g[v] * 4.0 - U2g[v]) / (2 * c.dtheta);
                            R[v] -= (-d * om) * V;
                        }
                    }
                }
                if (c.bfm_active && (fl & 2)) {
                    double S[5], fi[3], fv[3], dv;
                    bfm_source(c, s, Wc, V, S, fi, fv, dv);
#pragma unroll
                    for (int v = 0; v < 5; ++v) R[v] -= S[v];
                    if (st.last && (fl & 4)) {
                        c.finv[3 * n] = fi[0]; c.finv[3 * n + 1] = fi[1]; c.finv[3 * n + 2] = fi[2];
                        c.fvisc[3 * n] = fv[0]; c.fvisc[3 * n + 1] = fv[1]; c.fvisc[3 * n + 2] = fv[2];
                        c.dev[n] = dv;
                    }
                }
            }
            if (THREED && c.periodic && (k == 0 || k == NK - 1)) {
                double* rp = st.rper + (long)(k == 0 ? 0 : 5) * NI * NJ;
#pragma unroll
                for (int v = 0; v < 5; ++v) rp[(long)v * NI * NJ + (long)j * NI + i] = R[v];
            } else {
                const double dtv = (c.dt_mode == 1) ? c.dtmin[0] : c.dt[n];
                const double factor = st.alpha * dtv / c.vol[n];
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    st.Uout[v * N + n] = st.U0[v * N + n] - R[v] * factor;
                    nrm[v] += R[v] * R[v];
                }
            }
        }
#pragma unroll
        for (int v = 0; v < 5; ++v) FKl[v] = FKr[v];
        __syncthreads();
    }
    // ---- per-block norm partial sums (deterministic tree)
    if (st.last) {
        const long b = ((long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        for (int v = 0; v < 5; ++v) {
            L.red[tid] = nrm[v];
            __syncthreads();
            for (int s = NT / 2; s > 0; s >>= 1) {
                if (tid < s) L.red[tid] += L.red[tid + s];
                __syncthreads();
            }
            if (tid == 0) st.partial[b * 5 + v] = L.red[0];
            __syncthreads();
        }
    }
}

// Periodic closure of the k = 0 / k = NK-1 node-mates: residual averaging
// (solver.cpp:1836-1853), RK update of both, then solution averaging
// (solver.cpp:2308-2327).  Writes its own norm partials after the stage kernel's.
__global__ void __launch_bounds__(256) k_periodic(CaseDev c, StageDev st, long partial_offset) {
    __shared__ double red[256];
    if (*c.stop) return;
    const long NI = c.ni, NJ = c.nj, NK = c.nk, N = c.N;
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    double nrm[5] = {0, 0, 0, 0, 0};
    if (q < NI * NJ) {
        const long n0 = q, n1 = (NK - 1) * NI * NJ + q;
        double R1[5], R2[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            R1[v] = st.rper[(long)v * NI * NJ + q];
            R2[v] = st.rper[(long)(5 + v) * NI * NJ + q];
        }
        rotx(R2, c.per_cm, c.per_sm);
        double avg[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) avg[v] = (R1[v] + R2[v]) * 0.5;
        double back[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) back[v] = avg[v];
        rotx(back, c.per_c, c.per_s);
        const double dt0 = (c.dt_mode == 1) ? c.dtmin[0] : c.dt[n0];
        const double dt1 = (c.dt_mode == 1) ? c.dtmin[0] : c.dt[n1];
        const double f0 = st.alpha * dt0 / c.vol[n0];
        const double f1 = st.alpha * dt1 / c.vol[n1];
        double Ua[5], Ub[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            Ua[v] = st.U0[v * N + n0] - avg[v] * f0;
            Ub[v] = st.U0[v * N + n1] - back[v] * f1;
            nrm[v] = avg[v] * avg[v] + back[v] * back[v];
        }
        rotx(Ub, c.per_cm, c.per_sm);
        double Uavg[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) Uavg[v] = (Ua[v] + Ub[v]) * 0.5;
        double Uback[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) Uback[v] = Uavg[v];
        rotx(Uback, c.per_c, c.per_s);
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            st.Uout[v * N + n0] = Uavg[v];
            st.Uout[v * N + n1] = Uback[v];
        }
    }
    if (st.last) {
        for (int v = 0; v < 5; ++v) {
            red[threadIdx.x] = nrm[v];
            __syncthreads();
            for (int s = 128; s > 0; s >>= 1) {
                if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
                __syncthreads();
            }
            if (threadIdx.x == 0) st.partial[(partial_offset + blockIdx.x) * 5 + v] = red[0];
            __syncthreads();
        }
    }
}

// Residual norms of the last stage (solver.cpp:1488-1508), the baseline of
// iterations 1-5 (solver.cpp:1426-1439) and the convergence test
// (solver.cpp:2985-3015), all on the device.  Sets *stop once converged or diverged.
struct ConvDev {
    double* hist;        // [max_iter][5] log residuals
    double* init;        // [5] baseline
    long* conv_at;       // iteration of convergence, 0 if none
    int* diverged;       // 1 + variable index
    long hist_len;
    double drop;
    int steady;
};

__global__ void __launch_bounds__(256) k_norms(CaseDev c, const double* partial, int nparts, ConvDev cv) {
    __shared__ double red[256];
    __shared__ double lr[5];
    if (*c.stop) return;
    for (int v = 0; v < 5; ++v) {
        double s = 0.0;
        for (int q = threadIdx.x; q < nparts; q += 256) s += partial[(long)q * 5 + v];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            double nrm = sqrt(red[0]);
            double l;
            if (isnan(nrm) || isinf(nrm)) {
                l = nrm;
                if (*cv.diverged == 0) *cv.diverged = 1 + v;
            } else if (nrm >= 2.2250738585072014e-308) {
                l = log10(nrm / (double)((long)c.ni * c.nj * c.nk));
            } else {
                l = -16.0;
            }
            if (l < -16.0) l = -16.0;
            lr[v] = l;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const long it = *c.iter;
        if (it >= 1 && it <= cv.hist_len)
            for (int v = 0; v < 5; ++v) cv.hist[(it - 1) * 5 + v] = lr[v];
        if (it == 1) {
            for (int v = 0; v < 5; ++v) cv.init[v] = lr[v];
        } else if (it <= 5) {
            for (int v = 0; v < 5; ++v) cv.init[v] = dmaxd(cv.init[v], lr[v]);
        }
        if (*cv.diverged) {
            *c.stop = 1;
        } else if (cv.steady) {
            bool ok = true;
            for (int v = 0; v < 5 && ok; ++v) {
                if (v == 3 && c.topo != TOPO_3D && c.topo != TOPO_AXI) continue;
                if (v == 2 && c.topo == TOPO_1D) continue;
                if (cv.init[v] <= -15.99) continue;
                if (lr[v] <= -15.0) continue;
                if (lr[v] >= cv.init[v] - cv.drop) ok = false;
            }
            if (ok) {
                *cv.conv_at = it;
                *c.stop = 1;
            }
        }
        *c.iter = it + 1;
    }
}

// Per-iteration bookkeeping before the stages: keep the iteration counter.
__global__ void k_noop(CaseDev c) {}

}  // namespace dev
}  // namespace tbfm
