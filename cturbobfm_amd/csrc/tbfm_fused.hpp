// The fused residual + RK stage kernel.
//
// One 256-thread block owns a TI x TJ tile of (i, j) columns and marches along k over
// `kc` planes.  Per plane it stages the tile plus a 2-node halo in LDS (JST: conserved
// state, primitives, pressure and spectral factor |u| + a; Roe: primitives only).  The
// own node of each column comes from a per-thread register queue along k (planes
// k-1..k+2, which also feeds the K faces); the halo is filled by consecutive lanes, and
// ghost nodes at the domain ends are built exactly as the reference extrapolates them.
// Every I and J face of the tile is then evaluated once in a single divergence-free
// pass over a flat face index, with the face's unit normal and area read from
// host-precomputed records (the same IEEE operations the reference applies per call).
// Face results go straight to LDS arrays; the residual norms accumulate in registers
// and are reduced per block after the last plane.  The FP64 instances need about 240
// registers (two waves per SIMD); for JST on the ideal gas with FP32 flux the staged
// plane and the column queue are float (StageST), which gives three.
// The residual of each node is assembled in the reference's order
//   (I left, I right, J left, J right, K left, K right, geometric, Gong, BFM)
// (solver.cpp:1794-1821, 1997-2006, 2853-2891) and the RK update
// U = U0 - R * (alpha dt / V) (solver.cpp:2230-2249) is applied.  Nodes on the periodic
// planes k = 0 and k = NK-1 park their residual for the periodic-closure kernel.
#pragma once
#include <type_traits>

#include "tbfm_stage.hpp"

namespace tbfm {
namespace dev {

// A tile of TI x TJ (i, j) columns, one per thread of the block (NC <= 256: the threads
// beyond the tile only stage halo positions and evaluate faces); the LDS plane adds a
// 2-node halo on every side.
template <int TI_, int TJ_>
struct Tile {
    static constexpr int TI = TI_, TJ = TJ_;
    static constexpr int NC = TI * TJ;
    static constexpr int HI = TI + 4, HJ = TJ + 4, NP = HI * HJ;
    static constexpr int NH = NP - NC;   // halo positions
    static constexpr int NFI = 5 * TJ * (TI + 1), NFJ = 5 * (TJ + 1) * TI;
    static_assert(NC <= NT && NC > NT - 64, "a tile fills all but part of one wavefront");
    static_assert(TJ * (TI + 1) + (TJ + 1) * TI <= 3 * NT, "at most three faces per thread");
};

// The staged plane in the storage type ST of the kernel instance (double; float for the
// FP32 JST kernel) and the face fluxes of the tile, always double (boundary fluxes come
// from the FP64 boundary buffer).  After the last plane the per-block norm tree runs in
// the plane's storage.
template <class TL, typename ST>
struct PlaneLds {
    ST U[5][TL::NP];
    struct {
        ST W[5][TL::NP];
        ST P[TL::NP];
        ST RF[TL::NP];
    } d;
    struct {
        double FI[TL::NFI];   // [v][tj][m]  m = 0..TI, I face of node i0+m (its left face)
        double FJ[TL::NFJ];   // [v][r][ti] r = 0..TJ
    } f;
    static_assert(sizeof(U) + sizeof(d) >= 5 * NT * sizeof(double), "the norm tree fits in the plane");
    __device__ __forceinline__ double (*red())[NT] { return reinterpret_cast<double (*)[NT]>(&U[0][0]); }
};

// Storage type of the staged states of a stage-kernel instance: float for the JST scheme
// on the ideal gas with FP32 flux arithmetic (its sensors and dissipation differences
// then run in FP32 too), double otherwise.
template <int SCH, typename FT, int FL>
using StageST = typename std::conditional<(SCH == 0 && FL == 0 && sizeof(FT) == 4), float, double>::type;
// Face-record type: the FP32 records of the FP32-flux kernels, the FP64 ones otherwise.
template <typename FT>
using RecT = typename std::conditional<(sizeof(FT) == 4), float, double>::type;

// Source term t (0..11) of slab node sl from the pre-stage buffer, in its stored type:
// FP64, or the float the FP32-source pre-stage stored for the FP32-flux kernels
// (FT = float; widened where it is used, so that the load need not be waited for early).
template <typename FT>
using SrcT = typename std::conditional<(sizeof(FT) == 4), float, double>::type;
template <typename FT>
__device__ __forceinline__ SrcT<FT> src_term(const CaseDev& c, int t, int sl) {
    if constexpr (sizeof(FT) == 4) return gld(reinterpret_cast<const float*>(c.srcbuf) + (long)t * c.ns, (unsigned)sl);
    else return gld(c.srcbuf + (long)t * c.ns, (unsigned)sl);
}

template <class TL>
__device__ __forceinline__ int fiI(int v, int tj, int m) { return (v * TL::TJ + tj) * (TL::TI + 1) + m; }
template <class TL>
__device__ __forceinline__ int fiJ(int v, int r, int ti) { return (v * (TL::TJ + 1) + r) * TL::TI + ti; }

// Reads the conserved state of LDS position `pos` of plane k: a real node or a ghost one
// step outside the domain in i or j (linear extrapolation, solver.cpp:1980-1993, formed
// in the storage type ST from the state rounded to it, as the static-schedule kernel
// extrapolates from its staged plane); false for a position no face reads.
template <class TL, typename ST>
__device__ __forceinline__ bool fetch_position(const CaseDev& c, const double* __restrict__ U, int pos, int i0,
                                               int j0, int k, double Un[5]) {
    const int hi = pos % TL::HI, hj = pos / TL::HI;
    const int gi = i0 - 2 + hi, gj = j0 - 2 + hj;
    const int NI = c.ni, NJ = c.nj;
    const int N = (int)c.N;
    const bool iin = (gi >= 0 && gi < NI), jin = (gj >= 0 && gj < NJ);
    if (iin && jin) {
        load5(U, N, nid(c, gi, gj, k), Un);
        return true;
    }
    const bool ighost = jin && (gi == -1 || gi == NI);
    const bool jghost = iin && (gj == -1 || gj == NJ) && c.topo != TOPO_1D;
    if (!ighost && !jghost) return false;
    // a = boundary node, b = its inner neighbour
    int na, nb;
    bool lo;
    if (ighost) {
        lo = gi < 0;
        na = nid(c, lo ? gi + 1 : gi - 1, gj, k);
        nb = nid(c, lo ? gi + 2 : gi - 2, gj, k);
    } else {
        lo = gj < 0;
        na = nid(c, gi, lo ? gj + 1 : gj - 1, k);
        nb = nid(c, gi, lo ? gj + 2 : gj - 2, k);
    }
    double A[5], B[5];
    load5(U, N, na, A);
    load5(U, N, nb, B);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const ST a = ST(A[v]), b = ST(B[v]);
        Un[v] = double(lo ? a - (b - a) : a + (a - b));
    }
    return true;
}

// Where LDS position `pos` of every plane comes from, resolved once per block: kind 0 a
// position no face reads, 1 a real node at in-plane offset a, 2 / 3 a ghost node built
// from boundary node a and its inner neighbour b (2: before the domain start, 3: after
// its end), exactly as fetch_position does plane by plane.
struct HaloSrc {
    int kind, a, b;
};
template <class TL>
__device__ __forceinline__ HaloSrc halo_src(const CaseDev& c, int pos, int i0, int j0) {
    const int hi = pos % TL::HI, hj = pos / TL::HI;
    const int gi = i0 - 2 + hi, gj = j0 - 2 + hj;
    const int NI = c.ni, NJ = c.nj;
    const bool iin = (gi >= 0 && gi < NI), jin = (gj >= 0 && gj < NJ);
    HaloSrc h{0, 0, 0};
    if (iin && jin) {
        h.kind = 1;
        h.a = gj * NI + gi;
        return h;
    }
    const bool ighost = jin && (gi == -1 || gi == NI);
    const bool jghost = iin && (gj == -1 || gj == NJ) && c.topo != TOPO_1D;
    if (!ighost && !jghost) return h;
    bool lo;
    if (ighost) {
        lo = gi < 0;
        h.a = gj * NI + (lo ? gi + 1 : gi - 1);
        h.b = gj * NI + (lo ? gi + 2 : gi - 2);
    } else {
        lo = gj < 0;
        h.a = (lo ? gj + 1 : gj - 1) * NI + gi;
        h.b = (lo ? gj + 2 : gj - 2) * NI + gi;
    }
    h.kind = lo ? 2 : 3;
    return h;
}
// The conserved state of a resolved position on the plane starting at node `plane`
// (ghosts extrapolated in the storage type ST, as fetch_position).
template <typename ST>
__device__ __forceinline__ bool fetch_src(const double* __restrict__ U, int N, const HaloSrc& h, int plane,
                                          double Un[5]) {
    if (h.kind == 0) return false;
    if (h.kind == 1) {
        load5(U, N, plane + h.a, Un);
        return true;
    }
    double A[5], B[5];
    load5(U, N, plane + h.a, A);
    load5(U, N, plane + h.b, B);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const ST a = ST(A[v]), b = ST(B[v]);
        Un[v] = double((h.kind == 2) ? a - (b - a) : a + (a - b));
    }
    return true;
}

// Stores what the faces read at LDS position `pos` (JST: state, primitives, pressure,
// spectral factor; Roe: primitives); unused positions get harmless constants.
template <class TL, int SCH, int FL, typename ST>
__device__ __forceinline__ void commit_position(const CaseDev& c, PlaneLds<TL, ST>& L, int pos, bool ok,
                                                const double Un[5]) {
    if (!ok) {
#pragma unroll
        for (int v = 0; v < 5; ++v) L.d.W[v][pos] = ST(1.0);
        if constexpr (SCH == 0) {
#pragma unroll
            for (int v = 0; v < 5; ++v) L.U[v][pos] = ST(1.0);
            L.d.P[pos] = ST(1.0);
            L.d.RF[pos] = ST(0.0);
        }
        return;
    }
    const ST Us[5] = {ST(Un[0]), ST(Un[1]), ST(Un[2]), ST(Un[3]), ST(Un[4])};
    ST W[5], p, rf;
    derive<ST, FL>(Us, W, p, rf, Gas<FL>(c));
#pragma unroll
    for (int v = 0; v < 5; ++v) L.d.W[v][pos] = W[v];
    if constexpr (SCH == 0) {
#pragma unroll
        for (int v = 0; v < 5; ++v) L.U[v][pos] = Us[v];
        L.d.P[pos] = p;
        L.d.RF[pos] = rf;
    }
}

template <class TL, int SCH, int FL, typename ST>
__device__ inline void fill_position(const CaseDev& c, const double* __restrict__ U, PlaneLds<TL, ST>& L, int pos,
                                     int i0, int j0, int k) {
    double Un[5];
    const bool ok = fetch_position<TL, ST>(c, U, pos, i0, j0, k, Un);
    commit_position<TL, SCH, FL, ST>(c, L, pos, ok, Un);
}

// Face flux*|S| between LDS positions l and r (with ll = l - st, rr = r + st) along i
// or j; n and area are the face's unit normal and |S|.
template <class TL, int SCH, typename FT, int FL, typename ST, typename RT>
__device__ __forceinline__ void lds_face(const CaseDev& c, const PlaneLds<TL, ST>& L, int l, int st, RT nx, RT ny,
                                         RT nz, RT area, double out[5]) {
    const int r = l + st, ll = l - st, rr = r + st;
    if constexpr (SCH == 0) {
        // JST sensors of the left and right nodes from the staged pressures
        const ST pll = L.d.P[ll], pl = L.d.P[l], pr = L.d.P[r], prr = L.d.P[rr];
        const ST sl = jst_sensor<ST>(pll, pl, pr);
        const ST sr = jst_sensor<ST>(pl, pr, prr);
        FT Wavg[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) Wavg[v] = (FT(L.d.W[v][l]) + FT(L.d.W[v][r])) * FT(0.5);
        FT rr_ = FT(smax(L.d.RF[l], L.d.RF[r]));
        FT s = FT(smax(sl, sr));
        FT psi2 = FT(0.5) * s * rr_;
        FT t4 = FT(0.02) * rr_ - FT(2.0) * psi2;
        FT psi4 = (FT(0.0) < t4) ? t4 : FT(0.0);
        FT F[5];
        phys_flux_n<FT, FL>(Wavg, FT(nx), FT(ny), FT(nz), Gas<FL>(c), F);
        const FT sm = FT(area);
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            const ST ul = L.U[v][l], ur = L.U[v][r];
            ST d1 = ur - ul;
            ST d3 = ((L.U[v][rr] - ur) - d1 * 2) + (ul - L.U[v][ll]);
            FT dis = FT(d1) * psi2 - FT(d3) * psi4;
            out[v] = double((F[v] - dis) * sm);
        }
    } else {
        static_assert(sizeof(ST) == 8, "the Roe faces read FP64 primitives");
        double Wll[5], Wl[5], Wr[5], Wrr[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            Wll[v] = L.d.W[v][ll]; Wl[v] = L.d.W[v][l]; Wr[v] = L.d.W[v][r]; Wrr[v] = L.d.W[v][rr];
        }
        roe_face<FT, FL>(Wll, Wl, Wr, Wrr, c.muscl, c.limiter, nx, ny, nz, area, Gas<FL>(c), out);
    }
}

// Column queue along k: at step k the slots 0..3 hold planes k-1, k, k+1, k+2.  JST
// keeps the conserved state and pressure of all four slots and the primitives and
// spectral factor of slots 1 and 2 only (those of a new plane are derived when it
// reaches slot 2); Roe keeps the primitives of all four.  Unused slot values (planes
// beyond -1 .. NK) are never read by a face.
template <int SCH, int FL, typename ST>
struct KQueue;

template <int FL, typename ST>
struct KQueue<0, FL, ST> {
    // primitives without the density (which is U[slot][0]) and spectral factor of slots
    // 1, 2, 3 (Wv[s - 1], RF[s - 1]): each plane is derived once, when it enters slot 3;
    // all in the storage type ST
    ST U[4][5], P[4], Wv[3][4], RF[3];
    __device__ __forceinline__ void put(const CaseDev& c, int s, const double* u) {
        ST w[5], rf;
#pragma unroll
        for (int v = 0; v < 5; ++v) U[s][v] = ST(u[v]);
        derive<ST, FL>(U[s], w, P[s], rf, Gas<FL>(c));
        if (s >= 1) {
#pragma unroll
            for (int v = 0; v < 4; ++v) Wv[s - 1][v] = w[v + 1];
            RF[s - 1] = rf;
        }
    }
    __device__ __forceinline__ void init_end(const CaseDev&) {}
    __device__ __forceinline__ void shift() {
#pragma unroll
        for (int s = 0; s < 3; ++s) {
#pragma unroll
            for (int v = 0; v < 5; ++v) U[s][v] = U[s + 1][v];
            P[s] = P[s + 1];
        }
#pragma unroll
        for (int s = 0; s < 2; ++s) {
#pragma unroll
            for (int v = 0; v < 4; ++v) Wv[s][v] = Wv[s + 1][v];
            RF[s] = RF[s + 1];
        }
    }
    __device__ __forceinline__ void after_shift(const CaseDev&) {}
    __device__ __forceinline__ void Wc(ST w[5]) const {
        w[0] = U[1][0];
#pragma unroll
        for (int v = 0; v < 4; ++v) w[v + 1] = Wv[0][v];
    }
    template <typename FT, typename RT>
    __device__ __forceinline__ void face(const CaseDev& c, const RT nn[4], double out[5]) const {
        const ST sl = jst_sensor<ST>(P[0], P[1], P[2]);
        const ST sr = jst_sensor<ST>(P[1], P[2], P[3]);
        const ST Wl[5] = {U[1][0], Wv[0][0], Wv[0][1], Wv[0][2], Wv[0][3]};
        const ST Wr[5] = {U[2][0], Wv[1][0], Wv[1][1], Wv[1][2], Wv[1][3]};
        jst_face<FT, FL>(U[0], U[1], U[2], U[3], Wl, Wr, RF[0], RF[1], sl, sr, nn[0], nn[1], nn[2], nn[3],
                         Gas<FL>(c), out);
    }
};

// The table fluid keeps only slots 1 and 2 derived (its derivations cost registers that
// would take the kernel to one wavefront per SIMD): a plane is derived for its pressure
// when it enters slot 3 and again for its primitives when it reaches slot 2.
template <>
struct KQueue<0, 1, double> {
    static constexpr int FL = 1;
    double U[4][5], P[4], Wv[2][4], RF[2];
    __device__ __forceinline__ void put(const CaseDev& c, int s, const double* u) {
        double w[5], rf;
#pragma unroll
        for (int v = 0; v < 5; ++v) U[s][v] = u[v];
        derive<double, FL>(U[s], w, P[s], rf, Gas<FL>(c));
    }
    __device__ __forceinline__ void prim(const CaseDev& c, int s, int d) {   // primitives of slot s into d
        double w[5], p;
        derive<double, FL>(U[s], w, p, RF[d], Gas<FL>(c));
#pragma unroll
        for (int v = 0; v < 4; ++v) Wv[d][v] = w[v + 1];
    }
    __device__ __forceinline__ void init_end(const CaseDev& c) {
        prim(c, 1, 0);
        prim(c, 2, 1);
    }
    __device__ __forceinline__ void shift() {
#pragma unroll
        for (int s = 0; s < 3; ++s) {
#pragma unroll
            for (int v = 0; v < 5; ++v) U[s][v] = U[s + 1][v];
            P[s] = P[s + 1];
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) Wv[0][v] = Wv[1][v];
        RF[0] = RF[1];
    }
    __device__ __forceinline__ void after_shift(const CaseDev& c) { prim(c, 2, 1); }
    __device__ __forceinline__ void Wc(double w[5]) const {
        w[0] = U[1][0];
#pragma unroll
        for (int v = 0; v < 4; ++v) w[v + 1] = Wv[0][v];
    }
    template <typename FT, typename RT>
    __device__ __forceinline__ void face(const CaseDev& c, const RT nn[4], double out[5]) const {
        const double sl = jst_sensor<double>(P[0], P[1], P[2]);
        const double sr = jst_sensor<double>(P[1], P[2], P[3]);
        const double Wl[5] = {U[1][0], Wv[0][0], Wv[0][1], Wv[0][2], Wv[0][3]};
        const double Wr[5] = {U[2][0], Wv[1][0], Wv[1][1], Wv[1][2], Wv[1][3]};
        jst_face<FT, FL>(U[0], U[1], U[2], U[3], Wl, Wr, RF[0], RF[1], sl, sr, nn[0], nn[1], nn[2], nn[3],
                         Gas<FL>(c), out);
    }
};

template <int FL, typename ST>
struct KQueue<1, FL, ST> {
    static_assert(sizeof(ST) == 8, "the Roe column queue holds FP64 primitives");
    double W[4][5];
    __device__ __forceinline__ void put(const CaseDev& c, int s, const double* u) {
        double uu[5], p, rf;
#pragma unroll
        for (int v = 0; v < 5; ++v) uu[v] = u[v];
        derive<double, FL>(uu, W[s], p, rf, Gas<FL>(c));
    }
    __device__ __forceinline__ void init_end(const CaseDev&) {}
    __device__ __forceinline__ void shift() {
#pragma unroll
        for (int s = 0; s < 3; ++s)
#pragma unroll
            for (int v = 0; v < 5; ++v) W[s][v] = W[s + 1][v];
    }
    __device__ __forceinline__ void after_shift(const CaseDev&) {}
    __device__ __forceinline__ void Wc(double w[5]) const {
#pragma unroll
        for (int v = 0; v < 5; ++v) w[v] = W[1][v];
    }
    template <typename FT, typename RT>
    __device__ __forceinline__ void face(const CaseDev& c, const RT nn[4], double out[5]) const {
        roe_face<FT, FL>(W[0], W[1], W[2], W[3], c.muscl, c.limiter, nn[0], nn[1], nn[2], nn[3], Gas<FL>(c), out);
    }
};

// Face fq of a tile: I faces first (row r, face m at i-index i0 + m), then J faces
// (face row r at j-index j0 + r, column m); bnd marks a domain-boundary face (lo: the
// start side) whose flux comes from the boundary buffer.
template <class TL>
struct FaceIdx {
    bool isI, valid, bnd, lo;
    int r, m, gi, gj;
    __device__ __forceinline__ FaceIdx(int fq, int i0, int j0, int NI, int NJ) {
        constexpr int TI = TL::TI, NFI = TL::TJ * (TL::TI + 1);
        isI = fq < NFI;
        const int qq = isI ? fq : fq - NFI;
        r = isI ? qq / (TI + 1) : qq / TI;
        m = isI ? qq % (TI + 1) : qq % TI;
        gi = i0 + m;
        gj = j0 + r;
        valid = isI ? (gj < NJ && gi <= NI) : (gi < NI && gj <= NJ);
        const int gf = isI ? gi : gj, nf = isI ? NI : NJ;
        lo = gf == 0;
        bnd = lo || gf == nf;
    }
};

// Unit normal and area of K face f (index (k*NJ + j)*NI + i: the face below plane k).
// Face record `rec` for the stage kernel of flux type FT: the FP64 records, or for the
// FP32-flux kernels the FP32 records (their face arithmetic takes the float value of
// every component, so the results are the same and the loads half as wide).
template <typename FT>
__device__ __forceinline__ void face_record(const CaseDev& c, unsigned rec, RecT<FT> nn[4]) {
    if constexpr (sizeof(FT) == 4) gld_face32(c.nrm32, rec, nn);
    else gld_face(c.nrm, rec, nn);
}

template <typename FT>
__device__ __forceinline__ void kface_normal(const CaseDev& c, int f, RecT<FT> nn[4]) {
    face_record<FT>(c, (unsigned)(c.noffK + f), nn);
}

// Unit normal and area of an interior I or J face on plane k.
template <class TL, typename FT>
__device__ __forceinline__ void face_normal(const CaseDev& c, const FaceIdx<TL>& f, int k, RecT<FT> nn[4]) {
    const int NI = c.ni, NJ = c.nj;
    const int fo = f.isI ? (k * NJ + f.gj) * (NI + 1) + f.gi : c.noffJ + (k * (NJ + 1) + f.gj) * NI + f.gi;
    face_record<FT>(c, (unsigned)fo, nn);
}

// LDS position of halo slot h (0 .. NP-NT-1): two full rows above, two below, and the two
// columns on each side of the tile rows.  Consecutive lanes take consecutive slots so the
// halo work occupies as few wavefronts as possible.
template <class TL>
__device__ __forceinline__ int halo_pos(int h) {
    constexpr int HI = TL::HI, TJ = TL::TJ;
    if (h < 2 * HI) return h;                                   // rows 0, 1
    h -= 2 * HI;
    if (h < 2 * HI) return (TJ + 2) * HI + h;                  // rows TJ+2, TJ+3
    h -= 2 * HI;
    const int row = 2 + h / 4, col = h % 4;                     // side columns of tile rows
    return row * HI + ((col < 2) ? col : TL::TI + col);
}

// Face s (0, 1, 2) of thread tid in the flat I/J face index of a tile: tid, tid + 256,
// and for the faces beyond 512 the threads from the top down.
__device__ __forceinline__ int face_of(int s, int tid) {
    return s == 0 ? tid : (s == 1 ? tid + NT : 3 * NT - 1 - tid);
}

// The stage operands of one block: its case's buffers for this launch's roles.
struct StageOps {
    const double* Uin;     // U^(k-1)
    const double* U0;      // U^(0)
    double* Uout;          // U^(k)
    double alpha;
    int last, rmode;
};

template <class TL, int SCH, bool THREED, typename FT, int FL>
__device__ __forceinline__ void stage_block(const CaseDev* __restrict__ cp, const Work w, const StageOps st) {
    constexpr int TI = TL::TI, TJ = TL::TJ, HI = TL::HI, NH = TL::NH;
    constexpr int NFI = TJ * (TI + 1), NFJ = (TJ + 1) * TI;
    using ST = StageST<SCH, FT, FL>;
    using RT = RecT<FT>;
    __shared__ PlaneLds<TL, ST> L;
    const CaseDev& c = *cp;
    const int tid = threadIdx.x;
    const int ti = tid % TI, tj = tid / TI;
    const int i0 = w.x * TI, j0 = w.y * TJ;
    const int i = i0 + ti, j = j0 + tj;
    const int NI = c.ni, NJ = c.nj, NK = c.nk;
    const bool col = tid < TL::NC;   // the thread owns a column of the tile
    const bool act = col && (i < NI) && (j < NJ);
    // the block's k planes: piece w.z of sgz balanced pieces of [0, NK)
    const int kc0 = (w.z * NK) / c.sgz;
    const int kc1 = ((w.z + 1) * NK) / c.sgz;
    const bool doJ = (c.topo != TOPO_1D);
    const int N = (int)c.N;
    const int NIJ = NI * NJ;
    const int own = (tj + 2) * HI + (ti + 2);
    const int nfaces = NFI + (doJ ? NFJ : 0);
    const bool bladecol = THREED && c.bfm_active && act && gld(c.blade2d, (unsigned)(j * NI + i));
    const bool srcOn = (c.topo == TOPO_AXI) || c.bfm_active;
    const double* __restrict__ Uin = st.Uin;

    // K queue of the column (planes k-1 .. k+2 at the top of step k)
    KQueue<SCH, FL, ST> q;
    double FKl[5] = {0, 0, 0, 0, 0};
    if (THREED && act) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int gk = kc0 - 2 + s;
            double u[5] = {1.0, 1.0, 1.0, 1.0, 1.0};
            if (gk >= -1 && gk <= NK) load_k(c, Uin, i, j, gk, u);
            q.put(c, s, u);
        }
        q.init_end(c);
        if (!bladecol) {
            if (kc0 == 0) {
#pragma unroll
                for (int v = 0; v < 5; ++v) FKl[v] = gld(c.bnd + c.boffK, (unsigned)((j * NI + i) * 6 + v));
            } else {
                RT nk[4];
                kface_normal<FT>(c, (kc0 * NJ + j) * NI + i, nk);
                q.template face<FT>(c, nk, FKl);
            }
        }
    }
    double nsq[5] = {0.0, 0.0, 0.0, 0.0, 0.0};   // residual-norm accumulators of the column
    // Software pipeline: the next step's column plane and halo slot are read while the
    // current plane's faces are evaluated.
    double pre[5];
    if (THREED && act && kc0 + 2 <= NK) load_k(c, Uin, i, j, kc0 + 2, pre);
    const bool hown = tid < NH;
    const int hpos = hown ? halo_pos<TL>(tid) : 0;
    const HaloSrc hsrc = hown ? halo_src<TL>(c, hpos, i0, j0) : HaloSrc{0, 0, 0};
    double hU[5];
    bool hok = false;
    if (hown) hok = fetch_src<ST>(Uin, N, hsrc, kc0 * NIJ, hU);

#pragma unroll 1
    for (int k = kc0; k < kc1; ++k) {
        const int n = (k * NJ + j) * NI + i;
        const bool parked = THREED && c.periodic && (k == 0 || k == NK - 1);
        // unit normal and area of the K face above the node and of the thread's first I/J
        // face of this plane, requested before the plane is staged
        RT nnk[4] = {RT(1.0), RT(0.0), RT(0.0), RT(1.0)};
        if (THREED && act && !bladecol && k + 1 < NK) kface_normal<FT>(c, ((k + 1) * NJ + j) * NI + i, nnk);
        RT nnx[4] = {RT(1.0), RT(0.0), RT(0.0), RT(1.0)};
        {
            const FaceIdx<TL> f(tid, i0, j0, NI, NJ);
            if (tid < nfaces && f.valid && !f.bnd) face_normal<TL, FT>(c, f, k, nnx);
        }
        // update operands, read early and consumed at the update
        double u0v[5], dtv = 0.0, volv = 1.0;
        if (act && !parked && !st.rmode) {
            load5(st.U0, N, n, u0v);
            dtv = (c.dt_mode == 1) ? gld(c.dtmin, 0u) : gld(c.dt, (unsigned)n);
            volv = gld(c.vol, (unsigned)n);
        }
        ST Wc[5];
        if (THREED) {
            if (act) {
                q.shift();
                if (k + 2 <= NK) q.put(c, 3, pre);
                q.after_shift(c);
                if (k + 1 < kc1 && k + 3 <= NK) load_k(c, Uin, i, j, k + 3, pre);
                q.Wc(Wc);
                // stage the own node (JST: state, primitives, pressure, spectral factor)
                if constexpr (SCH == 0) {
#pragma unroll
                    for (int v = 0; v < 5; ++v) { L.U[v][own] = q.U[1][v]; L.d.W[v][own] = Wc[v]; }
                    L.d.P[own] = q.P[1];
                    L.d.RF[own] = q.RF[0];
                } else {
#pragma unroll
                    for (int v = 0; v < 5; ++v) L.d.W[v][own] = Wc[v];
                }
            }
        } else if (act) {
            double Ud[5];
            load5(Uin, N, nid(c, i, j, k), Ud);
            const ST Uc[5] = {ST(Ud[0]), ST(Ud[1]), ST(Ud[2]), ST(Ud[3]), ST(Ud[4])};
            ST Pc, RFc;
            derive<ST, FL>(Uc, Wc, Pc, RFc, Gas<FL>(c));
#pragma unroll
            for (int v = 0; v < 5; ++v) L.d.W[v][own] = Wc[v];
            if constexpr (SCH == 0) {
#pragma unroll
                for (int v = 0; v < 5; ++v) L.U[v][own] = Uc[v];
                L.d.P[own] = Pc;
                L.d.RF[own] = RFc;
            }
        }
        // tile positions outside the domain (possibly ghosts) and the halo by contiguous lanes
        if (col && !act) fill_position<TL, SCH, FL, ST>(c, Uin, L, own, i0, j0, k);
        if (hown) {
            commit_position<TL, SCH, FL, ST>(c, L, hpos, hok, hU);
            if (k + 1 < kc1) hok = fetch_src<ST>(Uin, N, hsrc, (k + 1) * NIJ, hU);
        }
        for (int h = tid + NT; h < NH; h += NT) fill_position<TL, SCH, FL, ST>(c, Uin, L, halo_pos<TL>(h), i0, j0, k);
        __syncthreads();
        // ---- every I and J face of the tile once: fq < NFI are I faces (row r, face m),
        // then J faces (row r of faces, column m); one code path, no divergence.  Each
        // thread takes faces tid and tid + 256; the faces beyond 512 go to the last
        // threads (the halo slots went to the first ones), so that no wavefront does both
        // a halo position and a third face.  The unit normal and area of the thread's
        // next face are requested before the current face is evaluated (those of its
        // first face before the plane was staged).
#pragma unroll 1
        for (int s = 0; s < 3; ++s) {
            const int fq = face_of(s, tid);
            if (fq >= nfaces) continue;
            const FaceIdx<TL> f(fq, i0, j0, NI, NJ);
            const RT nn[4] = {nnx[0], nnx[1], nnx[2], nnx[3]};
            if (s < 2) {
                const int fn = face_of(s + 1, tid);
                if (fn < nfaces) {
                    const FaceIdx<TL> g(fn, i0, j0, NI, NJ);
                    if (g.valid && !g.bnd) face_normal<TL, FT>(c, g, k, nnx);
                }
            }
            if (!f.valid) continue;
            double F[5];
            if (f.bnd) {
                const int nsz = f.isI ? NJ * NK : NI * NK;
                const int bq = f.isI ? k * NJ + f.gj : k * NI + f.gi;
                const int off = (f.isI ? 0 : c.boffJ) + ((f.lo ? 0 : nsz) + bq) * 6;
#pragma unroll
                for (int v = 0; v < 5; ++v) F[v] = gld(c.bnd, (unsigned)(off + v));
            } else {
                const int l = f.isI ? (f.r + 2) * HI + (f.m + 1) : (f.r + 1) * HI + (f.m + 2);
                lds_face<TL, SCH, FT, FL, ST, RT>(c, L, l, f.isI ? 1 : HI, nn[0], nn[1], nn[2], nn[3], F);
            }
            double* out = f.isI ? L.f.FI : L.f.FJ;
            const int base = f.isI ? f.r * (TI + 1) + f.m : f.r * TI + f.m;
            const int vstride = f.isI ? NFI : NFJ;
#pragma unroll
            for (int v = 0; v < 5; ++v) out[v * vstride + base] = F[v];
        }
        // ---- K face on the right of the node (face index k+1) from the column queue
        double FKr[5] = {0, 0, 0, 0, 0};
        if (THREED && act && !bladecol) {
            if (k + 1 == NK) {
#pragma unroll
                for (int v = 0; v < 5; ++v)
                    FKr[v] = gld(c.bnd + c.boffK, (unsigned)((NIJ + j * NI + i) * 6 + v));
            } else {
                q.template face<FT>(c, nnk, FKr);
            }
        }
        __syncthreads();
        // ---- residual assembly, sources, RK update
        if (act) {
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
            if (c.visc) viscous_add(c, i, j, k, n, R);
            if (srcOn && i >= c.bi0 && i < c.bi1) {
                // geometric, Gong, body force, subtracted in that order (the geometric
                // source's components 0, 1 and 4 are +0: subtracting them changes nothing)
                const int sl = (k * NJ + j) * (c.bi1 - c.bi0) + (i - c.bi0);
                R[2] -= double(src_term<FT>(c, 0, sl));
                R[3] -= double(src_term<FT>(c, 1, sl));
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int v = 0; v < 5; ++v) R[v] -= double(src_term<FT>(c, 2 + 5 * t + v, sl));
            }
            // no-slip walls after the periodic closure (solver.cpp:1828-1831): here for
            // the nodes outside the periodic planes, in k_periodic for theirs
            if (c.visc && !parked && c.nsw[n]) R[1] = R[2] = R[3] = R[0] * 0.0;
            if (parked) {
                double* rp = c.rper + (k == 0 ? 0 : 5) * NIJ;
#pragma unroll
                for (int v = 0; v < 5; ++v) gst(rp + (long)v * NIJ, (unsigned)(j * NI + i), R[v]);
            } else if (st.rmode) {
#pragma unroll
                for (int v = 0; v < 5; ++v) gst(st.Uout + (long)v * N, (unsigned)n, R[v]);
            } else {
                const double factor = st.alpha * dtv / volv;
#pragma unroll
                for (int v = 0; v < 5; ++v) gst(st.Uout + (long)v * N, (unsigned)n, u0v[v] - R[v] * factor);
                if (st.last)
#pragma unroll
                    for (int v = 0; v < 5; ++v) nsq[v] += R[v] * R[v];
            }
        }
#pragma unroll
        for (int v = 0; v < 5; ++v) FKl[v] = FKr[v];
        // No barrier here: the next plane's staging writes the plane buffers, whose last
        // readers (this plane's faces) finished before the barrier above, and its faces
        // overwrite the face buffers only after the staging barrier, which every thread
        // reaches after its assembly.
    }
    // ---- per-block norm partial sums (deterministic tree), in the primitive buffer: its
    // last readers finished before the last face barrier
    if (st.last) {
        double (*red)[NT] = L.red();
#pragma unroll
        for (int v = 0; v < 5; ++v) red[v][tid] = nsq[v];
        __syncthreads();
        const long b = ((long)w.z * c.sgy + w.y) * c.sgx + w.x;
        for (int s = NT / 2; s > 0; s >>= 1) {
            if (tid < s)
#pragma unroll
                for (int v = 0; v < 5; ++v) red[v][tid] += red[v][tid + s];
            __syncthreads();
        }
        if (tid < 5) gst(c.partial, (unsigned)(b * 5 + tid), red[tid][0]);
    }
}

// ------------------------------------------------------------------ JST stage, 3-D
// The 3-D JST instances run the same per-node arithmetic as stage_block with a static
// load schedule.  Vector-memory loads return in issue order and one counter tracks them,
// so a wait for any load is a wait for every load issued before it; the counter's
// value can only be known where every path through the loop issues the same loads.
// Every load of the k loop is therefore issued by every lane (lanes without a use read a
// safe address), in one order per plane:
//   top:     the column's plane k+3 and the halo slot's plane k+1 (used one plane later)
//   faces:   the K record and the three face-slot records of plane k+1
//   end:     U^(0), dt and V of plane k+1
// so that each value arrives a plane before it is used and no wait covers a younger
// load.  The boundary fluxes of domain-boundary I/J faces are read through the face
// slots (48-byte records behind the face records), and ghost positions of the staged
// plane are extrapolated from the staged nodes in LDS (blocks at a domain end only, one
// more barrier) instead of being loaded.  Loads whose results are used at once (the
// virtual plane NK of a non-periodic case, the end-plane K boundary fluxes, blade-row
// sources, viscous terms) stay where they are used.  The FP32 instance (ideal gas) stages
// the plane, the column queue and the face arithmetic in float like stage_block's FP32
// path; both extrapolate the i/j ghost positions in float from the state rounded to float
// (here from the staged plane), so every tiling gives the same bits.
// A per-plane copy of a value that the compiler may not treat as loop-invariant: what is
// derived from it (LDS and record offsets) is recomputed each plane instead of being
// hoisted out of the k loop into registers that the plane's arithmetic needs.
template <typename T>
__device__ __forceinline__ T per_plane(T x) {
    asm volatile("" : "+v"(x));
    return x;
}
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// 40 bytes from the record base at a face slot's offset: a face record (32 or 16 bytes,
// and what follows it) or a boundary-flux record (5 doubles).
struct SlotData {
    u32x4 a, b;
    u32x2 e;
};
__device__ __forceinline__ SlotData slot_load(const char* fb, unsigned off) {
    const TBFM_GLOBAL char* p = (const TBFM_GLOBAL char*)fb + off;
    SlotData d;
    d.a = __builtin_nontemporal_load((const TBFM_GLOBAL u32x4*)p);
    d.b = __builtin_nontemporal_load((const TBFM_GLOBAL u32x4*)(p + 16));
    d.e = __builtin_nontemporal_load((const TBFM_GLOBAL u32x2*)(p + 32));
    return d;
}
template <typename RT>
__device__ __forceinline__ void slot_record(const SlotData& d, RT nn[4]) {
    if constexpr (sizeof(RT) == 8) {
        const dpair x = __builtin_bit_cast(dpair, d.a), y = __builtin_bit_cast(dpair, d.b);
        nn[0] = x.x; nn[1] = x.y; nn[2] = y.x; nn[3] = y.y;
    } else {
        const fquad x = __builtin_bit_cast(fquad, d.a);
        nn[0] = x.x; nn[1] = x.y; nn[2] = x.z; nn[3] = x.w;
    }
}
__device__ __forceinline__ void slot_flux(const SlotData& d, double F[5]) {
    const dpair x = __builtin_bit_cast(dpair, d.a), y = __builtin_bit_cast(dpair, d.b);
    F[0] = x.x; F[1] = x.y; F[2] = y.x; F[3] = y.y; F[4] = __builtin_bit_cast(double, d.e);
}
// The K record at byte offset off (FP64: two 16-byte loads; FP32: one).
template <typename RT>
__device__ __forceinline__ void krec_load(const char* fb, unsigned off, RT nn[4]) {
    if constexpr (sizeof(RT) == 8) gld_face((const double*)fb, off / 32u, nn);
    else gld_face32((const float*)fb, off / 16u, nn);
}

// Face s (0..2) of thread tid in a tile, resolved once per block: bit 0 valid, bit 1 a
// domain-boundary face, bit 2 an I face, bits 3-12 the LDS position of its left node,
// bits 13-23 its index in the face buffers; b0 the byte offset of its plane-0 data from
// the record base (face record, or boundary-flux record), step the bytes per plane.
struct FaceSlot {
    unsigned b0;
    int info;
    __device__ __forceinline__ bool valid() const { return info & 1; }
    __device__ __forceinline__ bool bnd() const { return info & 2; }
    __device__ __forceinline__ bool isI() const { return info & 4; }
    __device__ __forceinline__ int l() const { return (info >> 3) & 1023; }
    __device__ __forceinline__ int obase() const { return info >> 13; }
};
// Slot 0 of thread tid is I face tid, slot 1 J face tid (so that both have a compile-time
// direction), slot 2 one of the remaining I faces (index 256 on) or J faces, taken by the
// threads from the top down.
template <class TL, unsigned RS>
__device__ __forceinline__ FaceSlot face_slot(const CaseDev& c, int s, int tid, int i0, int j0) {
    constexpr int TI = TL::TI, HI = TL::HI, NFI = TL::TJ * (TI + 1), NFJ = (TL::TJ + 1) * TI;
    static_assert(TL::NP <= 1024 && NFI + NFJ <= 2048, "face-slot fields fit their bits");
    static_assert(NFI >= NT && NFJ >= NT && NFI + NFJ <= 3 * NT, "one I and one J face per thread, a third for some");
    const int NI = c.ni, NJ = c.nj, NK = c.nk;
    FaceSlot d{0u, 0};
    int fq;
    if (s == 0) {
        fq = tid;
    } else if (s == 1) {
        fq = NFI + tid;
    } else {
        const int e = NT - 1 - tid;
        fq = (e < NFI - NT) ? NT + e : NFI + NT + (e - (NFI - NT));
        if (e >= NFI + NFJ - 2 * NT) return d;
    }
    const FaceIdx<TL> f(fq, i0, j0, NI, NJ);
    if (!f.valid) return d;
    const int l = f.isI ? (f.r + 2) * HI + (f.m + 1) : (f.r + 1) * HI + (f.m + 2);
    const int ob = f.isI ? f.r * (TI + 1) + f.m : f.r * TI + f.m;
    d.info = 1 | (f.bnd ? 2 : 0) | (f.isI ? 4 : 0) | (l << 3) | (ob << 13);
    if (f.bnd) {
        // boundary-flux record (side, k*NJ + j) of the I block, (side, k*NI + i) of the J block
        const int side = f.lo ? 0 : 1;
        const unsigned q = f.isI ? (unsigned)(side * NJ * NK + f.gj) : (unsigned)(2 * NJ * NK + side * NI * NK + f.gi);
        d.b0 = c.bnd_off + 48u * q;
    } else {
        d.b0 = RS * (unsigned)(f.isI ? f.gj * (NI + 1) + f.gi : c.noffJ + f.gj * NI + f.gi);
    }
    return d;
}
// Byte offset of a face slot's data on plane k (record 0 for a slot without a face).
template <unsigned RS>
__device__ __forceinline__ unsigned slot_off(const CaseDev& c, const FaceSlot& d, int k) {
    const unsigned NI = c.ni, NJ = c.nj;
    const unsigned step = d.bnd() ? 48u * (d.isI() ? NJ : NI) : RS * (d.isI() ? NJ * (NI + 1) : (NJ + 1) * NI);
    return d.valid() ? d.b0 + (unsigned)k * step : 0u;
}

// LDS of the 3-D JST kernel: the staged plane (conserved state, velocity and specific
// total energy, pressure, spectral factor; the density is U[0]), the I/J face fluxes of
// the tile, the K flux of the face below each column's next node and the column's
// residual-norm accumulators (which also hold the per-block norm tree).  The staged plane
// is in the storage type ST (float for the FP32 JST kernel, which has registers to spare
// for its norm accumulators and runs its norm tree in the staged plane).
struct ColumnNorms {
    double nsq[5][NT];
};
struct NoColumnNorms {};
template <class TL, typename ST = double>
struct PlaneLdsJ : std::conditional<sizeof(ST) == 8, ColumnNorms, NoColumnNorms>::type {
    ST U[5][TL::NP];
    ST W[4][TL::NP];   // W[v - 1]: primitive v = 1..4
    ST P[TL::NP];
    ST RF[TL::NP];
    double FI[5 * TL::TJ * (TL::TI + 1)];
    double FJ[5 * (TL::TJ + 1) * TL::TI];
    double FK[5][NT];
    static_assert(sizeof(ST) == 8 || 11 * TL::NP * sizeof(ST) >= 5 * NT * sizeof(double),
                  "the norm tree fits in the staged plane");
    __device__ __forceinline__ double (*red())[NT] { return reinterpret_cast<double (*)[NT]>(&U[0][0]); }
};

// Stores what the faces read at LDS position pos (state, primitives, pressure, spectral
// factor); unused positions get harmless constants (commit_position for this layout).
// The state is given in the storage type (converted from FP64 by the caller).
template <class TL, int FL, typename ST>
__device__ __forceinline__ void commit_j(const CaseDev& c, PlaneLdsJ<TL, ST>& L, int pos, bool ok, const ST Un[5]) {
    if (!ok) {
#pragma unroll
        for (int v = 0; v < 5; ++v) L.U[v][pos] = ST(1.0);
#pragma unroll
        for (int v = 0; v < 4; ++v) L.W[v][pos] = ST(1.0);
        L.P[pos] = ST(1.0);
        L.RF[pos] = ST(0.0);
        return;
    }
    ST W[5], p, rf;
    derive<ST, FL>(Un, W, p, rf, Gas<FL>(c));
#pragma unroll
    for (int v = 0; v < 5; ++v) L.U[v][pos] = Un[v];
#pragma unroll
    for (int v = 0; v < 4; ++v) L.W[v][pos] = W[v + 1];
    L.P[pos] = p;
    L.RF[pos] = rf;
}

// JST face flux*|S| between LDS positions l and r = l + st (lds_face for this layout; st is
// the direction's LDS stride, 1 for I faces and the halo row length for J faces: SD when
// the caller's direction is fixed at compile time, else the argument sdr).  The arithmetic
// runs in the storage type ST (the flux type of the kernel), the result is FP64.
template <class TL, int FL, int SD, typename ST>
__device__ __forceinline__ void lds_face_j(const CaseDev& c, const PlaneLdsJ<TL, ST>& L, int l, ST nx, ST ny,
                                           ST nz, ST area, double out[5], int sdr = 0) {
    const int st = SD ? SD : sdr;
    const int r = l + st, ll = l - st, rr = r + st;
    const ST pll = L.P[ll], pl = L.P[l], pr = L.P[r], prr = L.P[rr];
    const ST sl = jst_sensor<ST>(pll, pl, pr);
    const ST sr = jst_sensor<ST>(pl, pr, prr);
    ST Wavg[5];
    Wavg[0] = (L.U[0][l] + L.U[0][r]) * ST(0.5);
#pragma unroll
    for (int v = 1; v < 5; ++v) Wavg[v] = (L.W[v - 1][l] + L.W[v - 1][r]) * ST(0.5);
    const ST rr_ = smax(L.RF[l], L.RF[r]);
    const ST s = smax(sl, sr);
    const ST psi2 = ST(0.5) * s * rr_;
    const ST t4 = ST(0.02) * rr_ - ST(2.0) * psi2;
    const ST psi4 = (ST(0.0) < t4) ? t4 : ST(0.0);
    ST F[5];
    phys_flux_n<ST, FL>(Wavg, nx, ny, nz, Gas<FL>(c), F);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const ST ul = L.U[v][l], ur = L.U[v][r];
        const ST d1 = ur - ul;
        const ST d3 = ((L.U[v][rr] - ur) - d1 * 2) + (ul - L.U[v][ll]);
        const ST dis = d1 * psi2 - d3 * psi4;
        out[v] = double((F[v] - dis) * area);
    }
}

// A ghost position of the staged plane (one step outside the domain in i or j), built
// from the staged boundary node and its inner neighbour exactly as fetch_src builds it
// from global memory (solver.cpp:1980-1993), in the storage type; other positions are
// left alone.
template <class TL, int FL, typename ST>
__device__ __forceinline__ void ghost_from_lds(const CaseDev& c, PlaneLdsJ<TL, ST>& L, int p, int i0, int j0) {
    const int hi = p % TL::HI, hj = p / TL::HI;
    const int gi = i0 - 2 + hi, gj = j0 - 2 + hj;
    const int NI = c.ni, NJ = c.nj;
    const bool iin = (gi >= 0 && gi < NI), jin = (gj >= 0 && gj < NJ);
    const bool ighost = jin && (gi == -1 || gi == NI);
    const bool jghost = iin && (gj == -1 || gj == NJ);
    if (!ighost && !jghost) return;
    const bool lo = ighost ? (gi < 0) : (gj < 0);
    const int sd = ighost ? 1 : TL::HI;
    const int a = lo ? p + sd : p - sd, b = lo ? p + 2 * sd : p - 2 * sd;
    ST Un[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        const ST A = L.U[v][a], B = L.U[v][b];
        Un[v] = lo ? A - (B - A) : A + (A - B);
    }
    commit_j<TL, FL, ST>(c, L, p, true, Un);
}

// Column queue of the 3-D JST kernel along k.  At the K face of plane k it holds plane k-1
// (state and pressure), k+1 and k+2 (state, pressure, velocity and specific total energy,
// spectral factor) in registers; plane k itself is the column's own node of the staged
// plane in LDS.  The ideal gas derives a plane when it enters at k+2; the table fluid
// derives its pressure then and its primitives when it reaches k+1.  All in the storage
// type ST of the staged plane.
template <int FL, typename ST>
struct KColJ {
    ST A[5];
    ST sl;   // pressure sensor of plane k (the right sensor of the previous K face)
    ST C[5], Cp, Cw[4], Crf;
    ST D[5], Dp, Dw[4], Drf;
    __device__ __forceinline__ void enter(const CaseDev& c, const double u[5]) {
        ST w[5], rf;
#pragma unroll
        for (int v = 0; v < 5; ++v) D[v] = ST(u[v]);
        derive<ST, FL>(D, w, Dp, rf, Gas<FL>(c));
        if constexpr (FL == 0) {
#pragma unroll
            for (int v = 0; v < 4; ++v) Dw[v] = w[v + 1];
            Drf = rf;
        }
    }
    // plane k+2 moves to k+1
    __device__ __forceinline__ void d_to_c(const CaseDev& c) {
#pragma unroll
        for (int v = 0; v < 5; ++v) C[v] = D[v];
        Cp = Dp;
        if constexpr (FL == 0) {
#pragma unroll
            for (int v = 0; v < 4; ++v) Cw[v] = Dw[v];
            Crf = Drf;
        } else {
            ST w[5], p;
            derive<ST, FL>(C, w, p, Crf, Gas<FL>(c));
#pragma unroll
            for (int v = 0; v < 4; ++v) Cw[v] = w[v + 1];
        }
    }
    // plane k-1 from the own node of the staged plane (before it is overwritten)
    template <class TL>
    __device__ __forceinline__ void own_to_a(const PlaneLdsJ<TL, ST>& L, int own) {
#pragma unroll
        for (int v = 0; v < 5; ++v) A[v] = L.U[v][own];
    }
    // plane k+1 becomes the own node of the staged plane
    template <class TL>
    __device__ __forceinline__ void c_to_own(PlaneLdsJ<TL, ST>& L, int own) const {
#pragma unroll
        for (int v = 0; v < 5; ++v) L.U[v][own] = C[v];
#pragma unroll
        for (int v = 0; v < 4; ++v) L.W[v][own] = Cw[v];
        L.P[own] = Cp;
        L.RF[own] = Crf;
    }
    // K face between plane k (own LDS node) and k+1 (KQueue::face with slot 1 from LDS),
    // with the left sensor sl0; returns the right sensor (the next plane's left one)
    template <class TL>
    __device__ __forceinline__ ST face(const CaseDev& c, const PlaneLdsJ<TL, ST>& L, int own, ST sl0,
                                       const ST nn[4], double out[5]) const {
        ST B[5], Bw[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) B[v] = L.U[v][own];
        Bw[0] = B[0];
#pragma unroll
        for (int v = 0; v < 4; ++v) Bw[v + 1] = L.W[v][own];
        const ST Bp = L.P[own], Brf = L.RF[own];
        const ST sr = jst_sensor<ST>(Bp, Cp, Dp);
        const ST Wr[5] = {C[0], Cw[0], Cw[1], Cw[2], Cw[3]};
        jst_face<ST, FL>(A, B, C, D, Bw, Wr, Brf, Crf, sl0, sr, nn[0], nn[1], nn[2], nn[3], Gas<FL>(c), out);
        return sr;
    }
};

template <class TL, int FL, typename ST>
__device__ __forceinline__ void stage_block_jst3d(const CaseDev* __restrict__ cp, const Work w, const StageOps st) {
    using FT = ST;   // flux arithmetic and face records in the storage type
    constexpr int TI = TL::TI, TJ = TL::TJ, HI = TL::HI, NH = TL::NH, NP = TL::NP;
    constexpr int NFI = TJ * (TI + 1), NFJ = (TJ + 1) * TI;
    constexpr unsigned RS = 4u * sizeof(ST);   // face-record bytes (4 components)
    static_assert(sizeof(PlaneLdsJ<TL, ST>) <= 80 * 1024, "two blocks per CU");
    constexpr bool kNsqLds = sizeof(ST) == 8;   // FP64: norm accumulators in LDS
    static_assert(NH <= NT, "one halo slot per thread");
    __shared__ PlaneLdsJ<TL, ST> L;
    const CaseDev& c = *cp;
    const int tid = threadIdx.x;
    const int ti = tid % TI, tj = tid / TI;
    const int i0 = w.x * TI, j0 = w.y * TJ;
    const int i = i0 + ti, j = j0 + tj;
    const int NI = c.ni, NJ = c.nj, NK = c.nk;
    const bool col = tid < TL::NC;
    const bool act = col && (i < NI) && (j < NJ);
    const int kc0 = (w.z * NK) / c.sgz;
    const int kc1 = ((w.z + 1) * NK) / c.sgz;
    const int N = (int)c.N;
    const int NIJ = NI * NJ;
    const int own = (tj + 2) * HI + (ti + 2);
    // the column, clamped into the grid: the address lanes without a column read
    const int colc = min(j, NJ - 1) * NI + min(i, NI - 1);
    const bool bladecol = c.bfm_active && act && gld(c.blade2d, (unsigned)colc);
    const bool kOn = act && !bladecol;   // the column evaluates its K faces
    const bool srcOn = (c.topo == TOPO_AXI) || c.bfm_active;
    const double* __restrict__ Uin = st.Uin;
    const char* fb = c.fbase;
    // block-uniform: the staged region (tile and 2-node halo) holds ghost positions
    const bool ghosts = (i0 == 0) || (NI <= i0 + TI + 1) || (j0 == 0) || (NJ <= j0 + TJ + 1);
    // halo slot of the thread (threads beyond NH and slots that are not real nodes read
    // plane offset h.a, a real node or 0)
    const bool hown = tid < NH;
    const int hpos = hown ? halo_pos<TL>(tid) : 0;
    const HaloSrc hsrc = hown ? halo_src<TL>(c, hpos, i0, j0) : HaloSrc{0, 0, 0};
    FaceSlot fs[3];
#pragma unroll
    for (int s = 0; s < 3; ++s) fs[s] = face_slot<TL, RS>(c, s, tid, i0, j0);
    // plane read for the column queue's slot 3 at plane gk: the virtual plane NK of a
    // periodic case is plane 1 (rotated when used); other planes beyond the grid read
    // plane NK-1 (the virtual plane of other cases is built when used)
    auto qplane = [&](int gk) { return gk < NK ? gk : ((c.periodic && gk == NK) ? 1 : NK - 1); };
    // K record of plane k (the face above node k), record 0 where the column has none
    auto krec_off = [&](int k) -> unsigned {
        return (kOn && k + 1 < NK) ? RS * (unsigned)(c.noffK + (k + 1) * NIJ + colc) : 0u;
    };
    const double* dtp = (c.dt_mode == 1) ? c.dtmin : c.dt;

    // ---- column queue warm-up (planes kc0-2 .. kc0+1; plane kc0-1 as the own node of the
    // staged plane) and the K face below plane kc0
    KColJ<FL, ST> q;
    double nsq[5] = {0.0, 0.0, 0.0, 0.0, 0.0};   // FP32 plane: the column's norm accumulators
    {
        double FKl[5] = {0, 0, 0, 0, 0};
        if (act) {
            double u[5] = {1.0, 1.0, 1.0, 1.0, 1.0};
            if (kc0 - 2 >= -1) load_k(c, Uin, i, j, kc0 - 2, u);
            ST Ap;   // pressure of plane kc0-2 (left sensor of the K face below plane kc0)
            {
                ST w[5], rf;
#pragma unroll
                for (int v = 0; v < 5; ++v) q.A[v] = ST(u[v]);
                derive<ST, FL>(q.A, w, Ap, rf, Gas<FL>(c));
            }
#pragma unroll
            for (int v = 0; v < 5; ++v) u[v] = 1.0;
            load_k(c, Uin, i, j, kc0 - 1, u);   // kc0 - 1 >= -1
            {
                const ST us[5] = {ST(u[0]), ST(u[1]), ST(u[2]), ST(u[3]), ST(u[4])};
                commit_j<TL, FL, ST>(c, L, own, true, us);
            }
            load_k(c, Uin, i, j, kc0, u);
            q.enter(c, u);
            q.d_to_c(c);
#pragma unroll
            for (int v = 0; v < 5; ++v) u[v] = 1.0;
            if (kc0 + 1 <= NK) load_k(c, Uin, i, j, kc0 + 1, u);
            q.enter(c, u);
            // left sensor of the first K face of the loop: plane kc0 (planes kc0-1 .. kc0+1)
            q.sl = jst_sensor<ST>(L.P[own], q.Cp, q.Dp);
            if (!bladecol) {
                if (kc0 == 0) {
#pragma unroll
                    for (int v = 0; v < 5; ++v) FKl[v] = gld(c.bnd + c.boffK, (unsigned)((j * NI + i) * 6 + v));
                } else {
                    ST nk[4];
                    kface_normal<FT>(c, (kc0 * NJ + j) * NI + i, nk);
                    q.face(c, L, own, jst_sensor<ST>(Ap, L.P[own], q.Cp), nk, FKl);
                }
            }
        }
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            L.FK[v][tid] = FKl[v];
            if constexpr (kNsqLds) L.nsq[v][tid] = 0.0;
        }
    }
    // ---- loads ahead of plane kc0, in the order the loop issues them
    double pre[5], hU[5];
    load5(Uin, N, qplane(kc0 + 2) * NIJ + colc, pre);
    load5(Uin, N, kc0 * NIJ + hsrc.a, hU);
    ST nnk[4];
    krec_load<ST>(fb, krec_off(kc0), nnk);
    SlotData sd[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) sd[s] = slot_load(fb, slot_off<RS>(c, fs[s], kc0));
    double u0v[5], dtv, volv;
    {
        const int nn = kc0 * NIJ + colc;
        load5(st.U0, N, nn, u0v);
        dtv = gld(dtp, (c.dt_mode == 1) ? 0u : (unsigned)nn);
        volv = gld(c.vol, (unsigned)nn);
    }

#pragma unroll 1
    for (int k = kc0; k < kc1; ++k) {
        const int n = (k * NJ + j) * NI + i;
        const bool parked = c.periodic && (k == 0 || k == NK - 1);
        // ---- column queue (pre holds plane k+2) and the own node of the staged plane
        if (act) {
            q.own_to_a(L, own);
            q.c_to_own(L, own);
            q.d_to_c(c);
            if (k + 2 <= NK) {
                if (k + 2 == NK) {
                    if (c.periodic) rotx(pre, c.per_c, c.per_s);
                    else load_k(c, Uin, i, j, NK, pre);
                }
                q.enter(c, pre);
            }
        } else if (col) {
            const ST dummy[5] = {ST(1.0), ST(1.0), ST(1.0), ST(1.0), ST(1.0)};
            commit_j<TL, FL, ST>(c, L, own, false, dummy);
        }
        if (hown) {
            const ST hs[5] = {ST(hU[0]), ST(hU[1]), ST(hU[2]), ST(hU[3]), ST(hU[4])};
            commit_j<TL, FL, ST>(c, L, hpos, hsrc.kind == 1, hs);
        }
        // ---- loads for the next plane: the column's plane k+3, the halo slot's plane k+1.
        // The last plane of the piece has no next one: its loads read what this plane read
        // (cache-resident; the same loads in the same order keep every wait exact)
        const int kq = (k + 1 < kc1) ? k + 1 : k;
        load5(Uin, N, qplane(kq + 2) * NIJ + colc, pre);
        load5(Uin, N, min(kq, NK - 1) * NIJ + hsrc.a, hU);
        __syncthreads();
        if (ghosts) {
            ghost_from_lds<TL, FL, ST>(c, L, tid, i0, j0);
            if (tid + NT < NP) ghost_from_lds<TL, FL, ST>(c, L, tid + NT, i0, j0);
            __syncthreads();
        }
        // ---- every I and J face of the tile once (face slots of the thread); the few
        // threads with a third face (the last wavefront) read its data during the first
        SlotData s2d;
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const FaceSlot f{per_plane(fs[s].b0), per_plane(fs[s].info)};
            if (s == 1 && fs[2].valid()) s2d = slot_load(fb, slot_off<RS>(c, FaceSlot{per_plane(fs[2].b0), per_plane(fs[2].info)}, k));
            if (!f.valid()) continue;
            const SlotData& d = (s == 2) ? s2d : sd[s];
            double F[5];
            if (f.bnd()) {
                slot_flux(d, F);
            } else {
                ST nn[4];
                slot_record<ST>(d, nn);
                if (s == 0) lds_face_j<TL, FL, 1, ST>(c, L, f.l(), nn[0], nn[1], nn[2], nn[3], F);
                else if (s == 1) lds_face_j<TL, FL, HI, ST>(c, L, f.l(), nn[0], nn[1], nn[2], nn[3], F);
                else lds_face_j<TL, FL, 0, ST>(c, L, f.l(), nn[0], nn[1], nn[2], nn[3], F, f.isI() ? 1 : HI);
            }
            const bool fI = (s == 0) || (s == 2 && f.isI());
            double* out = fI ? L.FI : L.FJ;
            const int vstride = fI ? NFI : NFJ;
#pragma unroll
            for (int v = 0; v < 5; ++v) out[v * vstride + f.obase()] = F[v];
        }
        // ---- K face on the right of the node (face index k+1) from the column queue
        double FKr[5] = {0, 0, 0, 0, 0};
        if (kOn) {
            if (k + 1 == NK) {
#pragma unroll
                for (int v = 0; v < 5; ++v)
                    FKr[v] = gld(c.bnd + c.boffK, (unsigned)((NIJ + j * NI + i) * 6 + v));
            } else {
                q.sl = q.face(c, L, own, q.sl, nnk, FKr);
            }
        }
        // ---- the blade-row sources of the node (geometric, Gong, body force), read here
        // so that they arrive during the barrier; then the loads for the next plane: K
        // record and face slots
        const bool slab = act && srcOn && i >= c.bi0 && i < c.bi1;
        SrcT<FT> S[12];
        if (slab) {
            const int sl = (k * NJ + j) * (c.bi1 - c.bi0) + (i - c.bi0);
#pragma unroll
            for (int t = 0; t < 12; ++t) S[t] = src_term<FT>(c, t, sl);
        }
        {
            const int kn = min(kq, NK - 1);
            krec_load<ST>(fb, krec_off(kq), nnk);
#pragma unroll
            for (int s = 0; s < 2; ++s) sd[s] = slot_load(fb, slot_off<RS>(c, FaceSlot{per_plane(fs[s].b0), per_plane(fs[s].info)}, kn));
        }
        __syncthreads();
        // ---- residual assembly, sources, RK update
        if (act) {
            double R[5];
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                double r = 0.0;
                const double fl = L.FI[fiI<TL>(v, tj, ti)], fr = L.FI[fiI<TL>(v, tj, ti + 1)];
                r = (i == 0) ? r + fl : r - fl;
                r = r + fr;
                const double gl = L.FJ[fiJ<TL>(v, tj, ti)], gr = L.FJ[fiJ<TL>(v, tj + 1, ti)];
                r = (j == 0) ? r + gl : r - gl;
                r = r + gr;
                if (!bladecol) {
                    const double fkl = L.FK[v][tid];
                    r = (k == 0) ? r + fkl : r - fkl;
                    r = r + FKr[v];
                }
                R[v] = r;
            }
            if (c.visc) viscous_add(c, i, j, k, n, R);
            if (slab) {
                // geometric (y, z components; the others are +0), Gong, body force,
                // subtracted in that order
                R[2] -= double(S[0]);
                R[3] -= double(S[1]);
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int v = 0; v < 5; ++v) R[v] -= double(S[2 + 5 * t + v]);
            }
            // no-slip walls after the periodic closure (solver.cpp:1828-1831): here for
            // the nodes outside the periodic planes, in k_periodic for theirs
            if (c.visc && !parked && c.nsw[n]) R[1] = R[2] = R[3] = R[0] * 0.0;
            if (parked) {
                double* rp = c.rper + (k == 0 ? 0 : 5) * NIJ;
#pragma unroll
                for (int v = 0; v < 5; ++v) gst(rp + (long)v * NIJ, (unsigned)(j * NI + i), R[v]);
            } else if (st.rmode) {
#pragma unroll
                for (int v = 0; v < 5; ++v) gst(st.Uout + (long)v * N, (unsigned)n, R[v]);
            } else {
                const double factor = st.alpha * dtv / volv;
#pragma unroll
                for (int v = 0; v < 5; ++v) gst(st.Uout + (long)v * N, (unsigned)n, u0v[v] - R[v] * factor);
                if (st.last) {
#pragma unroll
                    for (int v = 0; v < 5; ++v) {
                        if constexpr (kNsqLds) L.nsq[v][tid] += R[v] * R[v];
                        else nsq[v] += R[v] * R[v];
                    }
                }
            }
        }
#pragma unroll
        for (int v = 0; v < 5; ++v) L.FK[v][tid] = FKr[v];
        // ---- loads for the next plane: U^(0), dt and V
        {
            const int nn = min(kq, NK - 1) * NIJ + colc;
            load5(st.U0, N, nn, u0v);
            dtv = gld(dtp, (c.dt_mode == 1) ? 0u : (unsigned)nn);
            volv = gld(c.vol, (unsigned)nn);
        }
    }
    // ---- per-block norm partial sums: a deterministic tree over the accumulators
    if (st.last) {
        double (*red)[NT];
        if constexpr (kNsqLds) {
            red = L.nsq;
        } else {
            // the staged plane's last readers (this plane's faces) are past the barrier
            // before the assembly
            red = L.red();
#pragma unroll
            for (int v = 0; v < 5; ++v) red[v][tid] = nsq[v];
        }
        __syncthreads();
        const long b = ((long)w.z * c.sgy + w.y) * c.sgx + w.x;
        for (int s = NT / 2; s > 0; s >>= 1) {
            if (tid < s)
#pragma unroll
                for (int v = 0; v < 5; ++v) red[v][tid] += red[v][tid + s];
            __syncthreads();
        }
        if (tid < 5) gst(c.partial, (unsigned)(b * 5 + tid), red[tid][0]);
    }
}

// Work: one entry per (i, j) tile and k chunk of a case.  The case and work tables are
// also passed as restrict kernel parameters (read-only for the launch), which lets the
// compiler keep case fields in scalar registers or reload them freely.
// The 3-D JST instances with the static load schedule: FP64 (ideal gas or table fluid) and
// FP32 (ideal gas) where the staged plane fits two blocks per CU and its halo one
// position per thread.
template <class TL, int SCH, bool THREED, typename FT, int FL>
constexpr bool kJst3d = SCH == 0 && THREED && (sizeof(FT) == 8 || FL == 0) && TL::NH <= NT &&
                        sizeof(PlaneLdsJ<TL, StageST<SCH, FT, FL>>) <= 80 * 1024;
// Waves per SIMD the register allocation must allow: two for the 3-D JST kernels of the
// ideal gas with the static schedule (LDS allows two blocks per CU; a third wave would
// spill the loads in flight, which costs far more than the wave gains), three for the
// other FP32 JST kernels (their LDS plane, 41.6 KB for 28x9 tiles, fits three blocks per
// CU), otherwise no bound (the other FP64 kernels need about 240 registers: two waves).
template <class TL, int SCH, bool THREED, typename FT, int FL>
constexpr int kStageWaves = (kJst3d<TL, SCH, THREED, FT, FL> && FL == 0) ? 2
                            : (std::is_same<StageST<SCH, FT, FL>, float>::value ? 3 : 1);

template <class TL, int SCH, bool THREED, typename FT, int FL>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(kStageWaves<TL, SCH, THREED, FT, FL>)))
k_stage(const CaseDev* __restrict__ cases, const Work* __restrict__ work, Launch La) {
    const Work w = work[blockIdx.x];
    if (w.cs < 0) return;   // padding entry of the XCD-interleaved table
    const CaseDev* __restrict__ cp = cases + w.cs;
    if (*cp->stop) return;
    StageOps st;
    st.Uin = stage_input(*cp, La);
    st.U0 = cp->U[0];
    st.Uout = cp->U[La.out];
    st.alpha = cp->alpha[La.stage];
    st.last = La.last;
    st.rmode = La.rmode;
    if constexpr (kJst3d<TL, SCH, THREED, FT, FL>)
        stage_block_jst3d<TL, FL, StageST<SCH, FT, FL>>(cp, w, st);
    else stage_block<TL, SCH, THREED, FT, FL>(cp, w, st);
}

// Periodic closure of the k = 0 / k = NK-1 node-mates: residual averaging
// (solver.cpp:1836-1853), RK update of both, then solution averaging
// (solver.cpp:2308-2327).  Writes its own norm partials after the stage kernel's.
// With st.rmode the averaged residuals themselves are written (residual entry point).
// Work: one entry per 256 node-mate pairs of a periodic 3-D case.
__global__ void __launch_bounds__(256) k_periodic(Launch La) {
    __shared__ double red[5][256];
    const Work w = La.work[blockIdx.x];
    const CaseDev& c = La.cases[w.cs];
    if (*c.stop) return;
    StageOps st;
    st.U0 = c.U[0];
    st.Uout = c.U[La.out];
    st.alpha = c.alpha[La.stage];
    st.last = La.last;
    st.rmode = La.rmode;
    const long NI = c.ni, NJ = c.nj, NK = c.nk, N = c.N;
    const long q = (long)w.x * 256 + threadIdx.x;
    double nrm[5] = {0, 0, 0, 0, 0};
    if (q < NI * NJ) {
        const long n0 = q, n1 = (NK - 1) * NI * NJ + q;
        double R1[5], R2[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            R1[v] = c.rper[(long)v * NI * NJ + q];
            R2[v] = c.rper[(long)(5 + v) * NI * NJ + q];
        }
        rotx(R2, c.per_cm, c.per_sm);
        double avg[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) avg[v] = (R1[v] + R2[v]) * 0.5;
        double back[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) back[v] = avg[v];
        rotx(back, c.per_c, c.per_s);
        if (c.visc) {
            if (c.nsw[n0]) avg[1] = avg[2] = avg[3] = avg[0] * 0.0;
            if (c.nsw[n1]) back[1] = back[2] = back[3] = back[0] * 0.0;
        }
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
        // the five equations' trees side by side (per equation the same pairings)
#pragma unroll
        for (int v = 0; v < 5; ++v) red[v][threadIdx.x] = nrm[v];
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (threadIdx.x < s)
#pragma unroll
                for (int v = 0; v < 5; ++v) red[v][threadIdx.x] += red[v][threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x < 5) c.partial[((long)c.nparts_stage + w.x) * 5 + threadIdx.x] = red[threadIdx.x][0];
    }
}

// Residual norms of the last stage (solver.cpp:1488-1508), the baseline of
// iterations 1-5 (solver.cpp:1426-1439) and the convergence test
// (solver.cpp:2985-3015), all on the device.  Sets *stop once converged or diverged and
// advances the iteration counter.  Work: one entry per case.
__global__ void __launch_bounds__(256) k_norms(Launch La) {
    __shared__ double red[5][256];
    __shared__ double lr[5];
    const Work w = La.work[blockIdx.x];
    const CaseDev& c = La.cases[w.cs];
    if (*c.stop) return;
    const ConvDev& cv = c.conv;
    const double* partial = c.partial;
    const int nparts = c.nparts_stage + c.nparts_per;
    // one pass over the partials for all five equations (per equation the same order of
    // additions as five separate passes)
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int q = threadIdx.x; q < nparts; q += 256)
#pragma unroll
        for (int v = 0; v < 5; ++v) acc[v] += partial[(long)q * 5 + v];
    // the five equations' trees side by side: per equation the pairing order of five
    // separate trees, with one barrier per level
#pragma unroll
    for (int v = 0; v < 5; ++v) red[v][threadIdx.x] = acc[v];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h)
#pragma unroll
            for (int v = 0; v < 5; ++v) red[v][threadIdx.x] += red[v][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < 5) {
        const int v = threadIdx.x;
        const double nrm = sqrt(red[v][0]);
        double l;
        if (isnan(nrm) || isinf(nrm)) l = nrm;
        else if (nrm >= 2.2250738585072014e-308) l = log10(nrm / (double)((long)c.ni * c.nj * c.nk));
        else l = -16.0;
        if (l < -16.0) l = -16.0;
        lr[v] = l;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // the first equation whose norm is not finite (solver.cpp:1494-1496)
        for (int v = 0; v < 5; ++v)
            if ((isnan(lr[v]) || isinf(lr[v])) && *cv.diverged == 0) *cv.diverged = 1 + v;
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
                if (cv.density_only && v > 0) break;
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
