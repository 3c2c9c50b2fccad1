// The fused residual + RK stage kernel.
//
// One 256-thread block owns a TI x TJ tile of (i, j) columns and marches along k over
// `kc` planes.  Per plane it stages the conserved state, primitives, pressure and the
// JST spectral factor |u| + a of the tile plus a 2-node halo in LDS (ghost nodes at the
// domain ends are built exactly as the reference extrapolates them), evaluates every I
// and J face of the tile once (JST pressure sensors inline from the staged pressures),
// and takes the K faces from a per-thread register queue of the column (planes
// k-1..k+2).  Face results go straight to their own LDS arrays (no register staging
// across a barrier) and the residual norms accumulate in LDS, which keeps the register
// budget at two waves per SIMD.
// The residual of each node is assembled in the reference's order
//   (I left, I right, J left, J right, K left, K right, geometric, Gong, BFM)
// (solver.cpp:1794-1821, 1997-2006, 2853-2891) and the RK update
// U = U0 - R * (alpha dt / V) (solver.cpp:2230-2249) is applied.  Nodes on the periodic
// planes k = 0 and k = NK-1 park their residual for the periodic-closure kernel.
#pragma once
#include "tbfm_stage.hpp"

namespace tbfm {
namespace dev {

template <int TI_, int TJ_>
struct Tile {
    static constexpr int TI = TI_, TJ = TJ_;
    static constexpr int HI = TI + 4, HJ = TJ + 4, NP = HI * HJ;
    static constexpr int NFI = 5 * TJ * (TI + 1), NFJ = 5 * (TJ + 1) * TI;
    static_assert(TI * TJ == NT, "tiles have 256 columns");
    static_assert(7 * NP >= NFI + NFJ, "face buffers must fit in the primitive region");
};

template <class TL>
struct PlaneLds {
    double U[5][TL::NP];
    struct {
        double W[5][TL::NP];
        double P[TL::NP];
        double RF[TL::NP];
    } d;
    struct {
        double FI[TL::NFI];   // [v][tj][m]  m = 0..TI, I face of node i0+m (its left face)
        double FJ[TL::NFJ];   // [v][r][ti] r = 0..TJ
    } f;
    double red[5][NT];        // per-thread residual-norm accumulators
};

template <class TL>
__device__ __forceinline__ int fiI(int v, int tj, int m) { return (v * TL::TJ + tj) * (TL::TI + 1) + m; }
template <class TL>
__device__ __forceinline__ int fiJ(int v, int r, int ti) { return (v * (TL::TJ + 1) + r) * TL::TI + ti; }

// Fills LDS position `pos` of plane k: real node, ghost (one step outside the domain in
// i or j) or unused.
template <class TL>
__device__ inline void fill_position(const CaseDev& c, const double* __restrict__ U, PlaneLds<TL>& L, int pos,
                                     int i0, int j0, int k) {
    const int hi = pos % TL::HI, hj = pos / TL::HI;
    const int gi = i0 - 2 + hi, gj = j0 - 2 + hj;
    const int NI = c.ni, NJ = c.nj;
    const bool iin = (gi >= 0 && gi < NI), jin = (gj >= 0 && gj < NJ);
    double Un[5];
    if (iin && jin) {
        load5(U, c.N, nid(c, gi, gj, k), Un);
    } else if (jin && (gi == -1 || gi == NI)) {
        double a[5], b[5];
        const int ia = (gi < 0) ? 0 : NI - 1, ib = (gi < 0) ? 1 : NI - 2;
        load5(U, c.N, nid(c, ia, gj, k), a);
        load5(U, c.N, nid(c, ib, gj, k), b);
#pragma unroll
        for (int v = 0; v < 5; ++v) Un[v] = (gi < 0) ? a[v] - (b[v] - a[v]) : a[v] + (a[v] - b[v]);
    } else if (iin && (gj == -1 || gj == NJ) && c.topo != TOPO_1D) {
        double a[5], b[5];
        const int ja = (gj < 0) ? 0 : NJ - 1, jb = (gj < 0) ? 1 : NJ - 2;
        load5(U, c.N, nid(c, gi, ja, k), a);
        load5(U, c.N, nid(c, gi, jb, k), b);
#pragma unroll
        for (int v = 0; v < 5; ++v) Un[v] = (gj < 0) ? a[v] - (b[v] - a[v]) : a[v] + (a[v] - b[v]);
    } else {
#pragma unroll
        for (int v = 0; v < 5; ++v) { L.U[v][pos] = 1.0; L.d.W[v][pos] = 1.0; }
        L.d.P[pos] = 1.0;
        L.d.RF[pos] = 0.0;
        return;
    }
    double W[5], p, rf;
    derive<double>(Un, W, p, rf, c.gamma, c.gm1);
#pragma unroll
    for (int v = 0; v < 5; ++v) { L.U[v][pos] = Un[v]; L.d.W[v][pos] = W[v]; }
    L.d.P[pos] = p;
    L.d.RF[pos] = rf;
}

// Face flux*|S| between LDS positions l and r (with ll, rr) along i or j.
template <class TL, int SCH, typename FT>
__device__ __forceinline__ void lds_face(const CaseDev& c, const PlaneLds<TL>& L, int ll, int l, int r, int rr,
                                         double sx, double sy, double sz, double out[5]) {
    if (SCH == 0) {
        // JST sensors of the left and right nodes from the staged pressures
        const double pll = L.d.P[ll], pl = L.d.P[l], pr = L.d.P[r], prr = L.d.P[rr];
        const double sl = jst_sensor<double>(pll, pl, pr);
        const double sr = jst_sensor<double>(pl, pr, prr);
        FT Wavg[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) Wavg[v] = (FT(L.d.W[v][l]) + FT(L.d.W[v][r])) * FT(0.5);
        FT rr_ = FT(dmaxd(L.d.RF[l], L.d.RF[r]));
        FT s = FT(dmaxd(sl, sr));
        FT psi2 = FT(0.5) * s * rr_;
        FT t4 = FT(0.02) * rr_ - FT(2.0) * psi2;
        FT psi4 = (FT(0.0) < t4) ? t4 : FT(0.0);
        FT F[5];
        FT sm = phys_flux<FT>(Wavg, FT(sx), FT(sy), FT(sz), FT(c.gm1), F);
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            const double ul = L.U[v][l], ur = L.U[v][r];
            double d1 = ur - ul;
            double d3 = ((L.U[v][rr] - ur) - d1 * 2) + (ul - L.U[v][ll]);
            FT dis = FT(d1) * psi2 - FT(d3) * psi4;
            out[v] = double((F[v] - dis) * sm);
        }
    } else {
        double Wll[5], Wl[5], Wr[5], Wrr[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            Wll[v] = L.d.W[v][ll]; Wl[v] = L.d.W[v][l]; Wr[v] = L.d.W[v][r]; Wrr[v] = L.d.W[v][rr];
        }
        roe_face<FT>(Wll, Wl, Wr, Wrr, c.muscl, c.limiter, sx, sy, sz, c.gamma, out);
    }
}

// K face from the column queue (slots: 0 = left-left, 1 = left, 2 = right, 3 = right-right)
template <int SCH, typename FT>
__device__ __forceinline__ void queue_face(const CaseDev& c, const double qU[4][5], const double qW[4][5],
                                           const double qP[4], const double qRF[4], double sx, double sy,
                                           double sz, double out[5]) {
    if (SCH == 0) {
        double sl = jst_sensor<double>(qP[0], qP[1], qP[2]);
        double sr = jst_sensor<double>(qP[1], qP[2], qP[3]);
        jst_face<FT>(qU[0], qU[1], qU[2], qU[3], qW[1], qW[2], qRF[1], qRF[2], sl, sr, sx, sy, sz, c.gm1, out);
    } else {
        roe_face<FT>(qW[0], qW[1], qW[2], qW[3], c.muscl, c.limiter, sx, sy, sz, c.gamma, out);
    }
}

template <class TL, int SCH, bool THREED, typename FT>
__global__ void __launch_bounds__(NT) k_stage(CaseDev c, StageDev st) {
    constexpr int TI = TL::TI, TJ = TL::TJ, HI = TL::HI, NP = TL::NP;
    __shared__ PlaneLds<TL> L;
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
    const bool srcOn = (c.topo == TOPO_AXI) || c.bfm_active;
    const double* __restrict__ Uin = st.Uin;

    // K queue: slots 0..3 hold planes k-1, k, k+1, k+2 at the top of step k
    double qU[4][5], qW[4][5], qP[4], qRF[4];
    double FKl[5] = {0, 0, 0, 0, 0};
    if (THREED && act) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int gk = kc0 - 2 + s;
            if (gk >= -1 && gk <= NK) {
                load_k(c, Uin, i, j, gk, qU[s]);
                derive<double>(qU[s], qW[s], qP[s], qRF[s], c.gamma, c.gm1);
            } else {
#pragma unroll
                for (int v = 0; v < 5; ++v) { qU[s][v] = 1.0; qW[s][v] = 1.0; }
                qP[s] = 1.0; qRF[s] = 0.0;
            }
        }
        if (!bladecol) {
            if (kc0 == 0) {
                const long q = (long)j * NI + i;
#pragma unroll
                for (int v = 0; v < 5; ++v) FKl[v] = c.bK[(long)v * NI * NJ + q];
            } else {
                const long sidx = ((long)kc0 * NJ + j) * NI + i;
                queue_face<SCH, FT>(c, qU, qW, qP, qRF, c.skx[sidx], c.sky[sidx], c.skz[sidx], FKl);
            }
        }
    }
    if (st.last)
#pragma unroll
        for (int v = 0; v < 5; ++v) L.red[v][tid] = 0.0;

#pragma unroll 1
    for (int k = kc0; k < kc1; ++k) {
        double Uc[5], Wc[5], Pc = 0.0, RFc = 0.0;
        if (THREED) {
            if (act) {
#pragma unroll
                for (int s = 0; s < 3; ++s) {
#pragma unroll
                    for (int v = 0; v < 5; ++v) { qU[s][v] = qU[s + 1][v]; qW[s][v] = qW[s + 1][v]; }
                    qP[s] = qP[s + 1];
                    qRF[s] = qRF[s + 1];
                }
                if (k + 2 <= NK) {
                    load_k(c, Uin, i, j, k + 2, qU[3]);
                    derive<double>(qU[3], qW[3], qP[3], qRF[3], c.gamma, c.gm1);
                }
#pragma unroll
                for (int v = 0; v < 5; ++v) { Uc[v] = qU[1][v]; Wc[v] = qW[1][v]; }
                Pc = qP[1];
                RFc = qRF[1];
            }
        } else if (act) {
            load5(Uin, N, nid(c, i, j, k), Uc);
            derive<double>(Uc, Wc, Pc, RFc, c.gamma, c.gm1);
        }
        // ---- stage plane k: own node from registers, halo by all threads
        if (act) {
#pragma unroll
            for (int v = 0; v < 5; ++v) { L.U[v][own] = Uc[v]; L.d.W[v][own] = Wc[v]; }
            L.d.P[own] = Pc;
            L.d.RF[own] = RFc;
        }
#pragma unroll 1
        for (int pos = tid; pos < NP; pos += NT) {
            const int hi = pos % HI, hj = pos / HI;
            const bool inTile = (hi >= 2 && hi < TI + 2 && hj >= 2 && hj < TJ + 2);
            if (inTile && (i0 + hi - 2) < NI && (j0 + hj - 2) < NJ) continue;  // owner wrote it
            fill_position<TL>(c, Uin, L, pos, i0, j0, k);
        }
        __syncthreads();
        // ---- faces, written straight to LDS
        double FKr[5] = {0, 0, 0, 0, 0};
        if (act) {
            double F[5];
            if (i == 0) {
                const long q = (long)k * NJ + j;
#pragma unroll
                for (int v = 0; v < 5; ++v) F[v] = c.bI[(long)v * NJ * NK + q];
            } else {
                const long sidx = ((long)k * NJ + j) * (NI + 1) + i;
                lds_face<TL, SCH, FT>(c, L, own - 2, own - 1, own, own + 1, c.six[sidx], c.siy[sidx], c.siz[sidx], F);
            }
#pragma unroll
            for (int v = 0; v < 5; ++v) L.f.FI[fiI<TL>(v, tj, ti)] = F[v];
            if (ti == TI - 1 || i == NI - 1) {
                if (i + 1 == NI) {
                    const long q = (long)k * NJ + j;
#pragma unroll
                    for (int v = 0; v < 5; ++v) F[v] = c.bI[(long)(5 + v) * NJ * NK + q];
                } else {
                    const long sidx = ((long)k * NJ + j) * (NI + 1) + i + 1;
                    lds_face<TL, SCH, FT>(c, L, own - 1, own, own + 1, own + 2, c.six[sidx], c.siy[sidx], c.siz[sidx], F);
                }
#pragma unroll
                for (int v = 0; v < 5; ++v) L.f.FI[fiI<TL>(v, tj, ti + 1)] = F[v];
            }
            if (doJ) {
                if (j == 0) {
                    const long q = (long)k * NI + i;
#pragma unroll
                    for (int v = 0; v < 5; ++v) F[v] = c.bJ[(long)v * NI * NK + q];
                } else {
                    const long sidx = ((long)k * (NJ + 1) + j) * NI + i;
                    lds_face<TL, SCH, FT>(c, L, own - 2 * HI, own - HI, own, own + HI, c.sjx[sidx], c.sjy[sidx],
                                          c.sjz[sidx], F);
                }
#pragma unroll
                for (int v = 0; v < 5; ++v) L.f.FJ[fiJ<TL>(v, tj, ti)] = F[v];
                if (tj == TJ - 1 || j == NJ - 1) {
                    if (j + 1 == NJ) {
                        const long q = (long)k * NI + i;
#pragma unroll
                        for (int v = 0; v < 5; ++v) F[v] = c.bJ[(long)(5 + v) * NI * NK + q];
                    } else {
                        const long sidx = ((long)k * (NJ + 1) + j + 1) * NI + i;
                        lds_face<TL, SCH, FT>(c, L, own - HI, own, own + HI, own + 2 * HI, c.sjx[sidx], c.sjy[sidx],
                                              c.sjz[sidx], F);
                    }
#pragma unroll
                    for (int v = 0; v < 5; ++v) L.f.FJ[fiJ<TL>(v, tj + 1, ti)] = F[v];
                }
            }
            if (THREED && !bladecol) {
                if (k + 1 == NK) {
                    const long q = (long)j * NI + i;
#pragma unroll
                    for (int v = 0; v < 5; ++v) FKr[v] = c.bK[(long)(5 + v) * NI * NJ + q];
                } else {
                    const long sidx = ((long)(k + 1) * NJ + j) * NI + i;
                    queue_face<SCH, FT>(c, qU, qW, qP, qRF, c.skx[sidx], c.sky[sidx], c.skz[sidx], FKr);
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
                const double fl = L.f.FI[fiI<TL>(v, tj, ti)], fr = L.f.FI[fiI<TL>(v, tj, ti + 1)];
                r = (i == 0) ? r + fl : r - fl;
                r = r + fr;
                if (doJ) {
                    const double gl = L.f.FJ[fiJ<TL>(v, tj, ti)], gr = L.f.FJ[fiJ<TL>(v, tj + 1, ti)];
                    r = (j == 0) ? r + gl : r - gl;
                    r = r + gr;
                }
                if (THREED && !bladecol) {
                    r = (k == 0) ? r + FKl[v] : r - FKl[v];
                    r = r + FKr[v];
                }
                R[v] = r;
            }
            if (srcOn && i >= c.bi0 && i < c.bi1) node_sources<THREED>(st.cdev, Uin, st.last, i, j, k, n, Wc, R);
            if (THREED && c.periodic && (k == 0 || k == NK - 1)) {
                double* rp = st.rper + (long)(k == 0 ? 0 : 5) * NI * NJ;
#pragma unroll
                for (int v = 0; v < 5; ++v) rp[(long)v * NI * NJ + (long)j * NI + i] = R[v];
            } else if (st.rmode) {
#pragma unroll
                for (int v = 0; v < 5; ++v) st.Uout[v * N + n] = R[v];
            } else {
                const double dtv = (c.dt_mode == 1) ? c.dtmin[0] : c.dt[n];
                const double factor = st.alpha * dtv / c.vol[n];
#pragma unroll
                for (int v = 0; v < 5; ++v) st.Uout[v * N + n] = st.U0[v * N + n] - R[v] * factor;
                if (st.last)
#pragma unroll
                    for (int v = 0; v < 5; ++v) L.red[v][tid] += R[v] * R[v];
            }
        }
#pragma unroll
        for (int v = 0; v < 5; ++v) FKl[v] = FKr[v];
        __syncthreads();
    }
    // ---- per-block norm partial sums (deterministic tree, through the U region)
    if (st.last) {
        const long b = ((long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        for (int s = NT / 2; s > 0; s >>= 1) {
            if (tid < s)
#pragma unroll
                for (int v = 0; v < 5; ++v) L.red[v][tid] += L.red[v][tid + s];
            __syncthreads();
        }
        if (tid < 5) st.partial[b * 5 + tid] = L.red[tid][0];
    }
}

// Periodic closure of the k = 0 / k = NK-1 node-mates: residual averaging
// (solver.cpp:1836-1853), RK update of both, then solution averaging
// (solver.cpp:2308-2327).  Writes its own norm partials after the stage kernel's.
// With st.rmode the averaged residuals themselves are written (residual entry point).
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
        if (st.rmode) {
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                st.Uout[v * N + n0] = avg[v];
                st.Uout[v * N + n1] = back[v];
            }
        } else {
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

}  // namespace dev
}  // namespace tbfm
