// Device functions and kernels of the explicit residual path for gfx950.
//
// Device layout (i fastest): node n = (k*NJ + j)*NI + i; conserved state is five SoA
// planes of N doubles.  Face vectors are SoA per component:
//   S_I: (k*NJ + j)*(NI+1) + i     S_J: (k*(NJ+1) + j)*NI + i     S_K: (k*NJ + j)*NI + i
// Boundary-face fluxes (already multiplied by |S|) live in per-side buffers written by
// the boundary kernel before each stage.
//
// Arithmetic follows the reference operation by operation (the oracle header lists
// the file:line of each formula); the library is compiled with -ffp-contract=off so
// that FP64 results match the reference CPU build bit for bit wherever no libm
// transcendental is involved.  The flux type FT is double (plain FP64 path) or float
// (mixed precision: flux, dissipation and source arithmetic in FP32, state
// differences, RK accumulation, norms and convergence in FP64).
#pragma once
#include <hip/hip_runtime.h>

namespace tbfm {
namespace dev {

constexpr int NT = 256;         // threads per stage-kernel block (tile columns)

enum : int { PT_INLET = 0, PT_OUTLET = 1, PT_RADIAL = 2, PT_WALL = 3, PT_FAKE = 4 };
enum : int { TOPO_2D = 0, TOPO_3D = 1, TOPO_AXI = 2, TOPO_1D = 3 };

struct PatchDev {
    int type;
    int profile;        // radial-equilibrium profile index
    double v0, v1;      // inlet: p0, T0; outlet: p
};

constexpr int MAX_PATCH = 32;
constexpr int MAX_PROF = 4;

// Everything a kernel needs; passed by value as the kernel argument.
struct CaseDev {
    int ni, nj, nk;
    long N;
    int topo, scheme, muscl, limiter;
    int periodic;
    double gamma, gm1, rgas, cp;
    double per_c, per_s;          // cos/sin of +angle
    double per_cm, per_sm;        // cos/sin of -angle
    double cfl, fixed_dt;
    int dt_mode;                  // 0 local, 1 global, 2 fixed
    double ramp_iter, init_p;
    int bfm_active, bfm_model, blockage_active;
    double dtheta;                // Gong: 2*pi/(NK-1)
    // geometry
    const double *six, *siy, *siz, *sjx, *sjy, *sjz, *skx, *sky, *skz;
    // unit normals S/|S| and areas |S| of every face (same layout), computed on the host
    // with the reference's own operations (Vector3D::magnitude, operator/)
    const double *nix, *niy, *niz, *nia, *njx, *njy, *njz, *nja, *nkx, *nky, *nkz, *nka;
    const double *vol;
    const double *dux, *duy, *duz;     // dual nodes (NI+1)(NJ+1)(NK+1)
    // boundary faces: patch id and inlet direction per face; flux buffers
    const int *pI, *pJ, *pK;           // [2][...] patch index
    const double *dirI, *dirJ, *dirK;  // [2][3][...]
    double *bI, *bJ, *bK;              // [2][5][...] flux*|S|
    PatchDev patch[MAX_PATCH];
    // radial equilibrium profiles
    int nprof;
    int prof_i[MAX_PROF], prof_j0[MAX_PROF], prof_n[MAX_PROF];
    double* prof_p;                    // [MAX_PROF][NJ]
    const double* prof_r;              // [MAX_PROF][NJ]
    const double* prof_cs;             // [MAX_PROF][NJ] cos(theta) at (i, j, 0)
    const double* prof_sn;
    // body force slab: nodes with i in [bi0, bi1)
    int bi0, bi1;
    const unsigned char* blade2d;      // (j*NI + i): bladePresent(i,j,0) > 0
    const unsigned char* sflag;        // bit0 present>0, bit1 |grad b|>1e-10, bit2 numberBlades != 0
    const double *s_cs, *s_sn, *s_r, *s_om, *s_b, *s_gbx, *s_gby, *s_gbz;
    const double *s_nax, *s_nrad, *s_ntan, *s_pitch, *s_spitch, *s_sol, *s_beta0, *s_kp;
    const double* axi_gf;              // axisymmetric: |S_K(i,j,0)| * wedge angle, (j*NI + i)
    // side outputs
    double *finv, *fvisc, *dev;
    // time step
    double* dt;                        // N
    double* dtmin;                     // [1]
    // convergence / control on device
    int* stop;                         // non-zero: skip all work
    long* iter;                        // current iteration counter (1-based during an iteration)
};

struct StageDev {
    const double* Uin;     // U^(k-1)
    const double* U0;      // U^(0)
    double* Uout;          // U^(k)
    double alpha;
    int last;              // last stage: norms and side outputs
    double* rper;          // [2][5][NI*NJ] residuals of the periodic planes
    double* partial;       // [blocks][5] norm partial sums
    int kc;                // k planes per block
    const struct CaseDev* cdev;   // device-resident copy of the case description
    int rmode;             // write the residual instead of the update (residual entry point)
};

__device__ __forceinline__ double dmaxd(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double dmind(double a, double b) { return (b < a) ? b : a; }

}  // namespace dev
}  // namespace tbfm
