// Device functions and kernels of the explicit residual path for gfx950.
//
// Device layout (i fastest): node n = (k*NJ + j)*NI + i; conserved state is five SoA
// planes of N doubles.  Face vectors are SoA per component:
//   S_I: (k*NJ + j)*(NI+1) + i     S_J: (k*(NJ+1) + j)*NI + i     S_K: (k*NJ + j)*NI + i
// Boundary-face fluxes (already multiplied by |S|) live in per-side 48-byte records
// written by the boundary kernel before each stage, behind the face records.
//
// Arithmetic follows the reference operation by operation (the oracle header lists
// the file:line of each formula); the library is compiled with -ffp-contract=off so
// that FP64 results match the reference CPU build bit for bit wherever no libm
// transcendental is involved.  The flux type FT is double (plain FP64 path) or float
// (mixed precision: flux, dissipation and source arithmetic in FP32, state
// differences, RK accumulation, norms and convergence in FP64).  The fluid is the ideal
// gas or the look-up-table fluid (FLUID_MODEL = real, FP64 only).
#pragma once
#include <hip/hip_runtime.h>

#include "tbfm_fluid.hpp"

namespace tbfm {
namespace dev {

constexpr int NT = 256;         // threads per stage-kernel block (tile columns)

enum : int { PT_INLET = 0, PT_OUTLET = 1, PT_RADIAL = 2, PT_WALL = 3, PT_FAKE = 4, PT_TRANSPARENT = 5 };
enum : int { TOPO_2D = 0, TOPO_3D = 1, TOPO_AXI = 2, TOPO_1D = 3 };

struct PatchDev {
    int type;
    int profile;        // radial-equilibrium profile index
    double v0, v1;      // inlet: p0, T0; outlet: p
    const double* tab;  // ideal-gas inlet: bisection table (below), null otherwise
};

// Ideal-gas subsonic inlet: along the pressure bisection of boundary_base.cpp:51-68 the
// static temperature T(p) and speed V(p) of the inlet total state depend only on the
// patch's p0, T0 and on the visited midpoint, never on the interior state.  The first
// INLET_LEVELS levels of the bisection tree (node q has children 2q+1: p_low moves up,
// and 2q+2: p_high moves down) are tabulated once per patch: T at tab[q], V at
// tab[INLET_NODES + q], computed with the same operations as the on-the-fly values.
constexpr int INLET_LEVELS = 16;
constexpr int INLET_NODES = (1 << INLET_LEVELS) - 1;

constexpr int MAX_PATCH = 32;
constexpr int MAX_PROF = 4;

// Residual norms, baseline and convergence state of one case (solver.cpp:1426-1439,
// 1488-1508, 2985-3015).
struct ConvDev {
    double* hist;        // [hist_len][5] log residuals of iterations 1..hist_len
    double* init;        // [5] baseline (max over iterations 1-5)
    long* conv_at;       // iteration of convergence, 0 if none
    int* diverged;       // 1 + variable index
    long hist_len;
    double drop;         // orders of magnitude
    int steady;
    int density_only;    // 1: converged when the density residual alone has dropped
};

// Everything the kernels need about one case.  Kept in device memory (one entry per
// case of a batch) and read by the kernels through the launch's case table, so that a
// captured launch sequence sees later updates (history growth, tiling, stop flags).
struct CaseDev {
    int ni, nj, nk;
    long N;
    int topo, scheme, muscl, limiter;
    int periodic;
    double gamma, gm1, rgas, cp;
    int real;                     // FLUID_MODEL = real: thermodynamics from the tables in lut
    const LutSet* lut;            // device copy (device pointers), null for the ideal gas
    double per_c, per_s;          // cos/sin of +angle
    double per_cm, per_sm;        // cos/sin of -angle
    double cfl, fixed_dt;
    int dt_mode;                  // 0 local, 1 global, 2 fixed
    double ramp_iter, init_p;
    int bfm_active, bfm_model, blockage_active;
    double dtheta;                // Gong: 2*pi/(NK-1)
    // geometry
    const double *six, *siy, *siz, *sjx, *sjy, *sjz, *skx, *sky, *skz;
    // unit normals S/|S| and areas |S| of every face, computed on the host with the
    // reference's own operations (Vector3D::magnitude, operator/), in one allocation of
    // 32-byte records (nx, ny, nz, |S|), faces in the order of S per direction: the I
    // faces from record 0, the J faces from record noffJ, the K faces from record noffK
    int fp32;             // GPU_PRECISION = mixed: FP32 flux, dissipation and source arithmetic
    const double* nrm;
    const float* nrm32;   // FP32-flux cases: the same records rounded to float (16 bytes each)
    int nfI, nfJ, nfK, noffJ, noffK;
    const double *vol;
    const double *dux, *duy, *duz;     // dual nodes (NI+1)(NJ+1)(NK+1)
    // boundary faces: patch id and inlet direction per face; flux buffers
    const int *pI, *pJ, *pK;           // [2][...] patch index
    const double *dirI, *dirJ, *dirK;  // [2][3][...]
    // boundary-face fluxes * |S|: per direction [2 sides][faces of a side][6] (5 doubles and
    // a pad, 48-byte records; bszI = NJ*NK, bszJ = NI*NK, bszK = NI*NJ), the I block at 0,
    // J at boffJ, K at boffK (doubles).  They live in the face-record allocation, bnd_off
    // bytes from its start (fbase: nrm or nrm32), so that the stage kernel reads a face's
    // record or its boundary flux from one base address.
    double* bnd;
    int boffJ, boffK;
    const char* fbase;
    unsigned bnd_off;
    PatchDev patch[MAX_PATCH];
    // radial equilibrium profiles
    int nprof;
    int prof_i[MAX_PROF], prof_j0[MAX_PROF], prof_n[MAX_PROF];
    // hub pressure of each outlet profile: 0 radial equilibrium (kept), 1 throttle
    // (par: k_t, inlet p0), 2 mass flow (par: target, relaxation, maximum change, initial
    // pressure; mfb: which turbo boundary carries the patch); radial: integrated radially,
    // else uniform.  need_mf: some profile reads the boundary mass flows.
    int prof_kind[MAX_PROF], prof_radial[MAX_PROF], prof_mfb[MAX_PROF], prof_need_mf;
    double prof_par[MAX_PROF][4];
    // Greitzer plenum (ENABLE_GREITZER_MODELING): every profile's hub is the plenum
    // pressure; gz = {plenum pressure, outlet mass flow, temperature}
    int greitzer;
    double gz_dt, gz_vol, gz_amb, gz_kt;
    double* gz;
    double* prof_p;                    // [MAX_PROF][NJ]
    const double* prof_r;              // [MAX_PROF][NJ]
    const double* prof_cs;             // [MAX_PROF][NJ] cos(theta) at (i, j, 0)
    const double* prof_sn;
    // body force slab: nodes with i in [bi0, bi1); ns = slab nodes; srcbuf holds the
    // source terms of the current stage, [12][ns]: the geometric source's y and z
    // components (its other three are +0), the Gong source, the body force; doubles, or
    // floats for the FP32-flux kernels (fp32: their sources are FP32 arithmetic)
    int bi0, bi1, ns;
    double* srcbuf;
    const unsigned char* blade2d;      // (j*NI + i): bladePresent(i,j,0) > 0
    const unsigned char* sflag;        // bit0 present>0, bit1 |grad b|>1e-10, bit2 numberBlades != 0
    const double *s_cs, *s_sn, *s_r, *s_om, *s_b, *s_gbx, *s_gby, *s_gbz;
    const double *s_nax, *s_nrad, *s_ntan, *s_pitch, *s_spitch, *s_sol, *s_beta0, *s_kp;
    // Hall-Thollet: streamwise length with the reference's leading-edge fallback,
    // deviation pivot (off-design) and stall deviation (STALLED_BFM_ACTIVE), per slab node
    const double *s_stwl, *s_piv, *s_stall;
    // PERTURBATION_BODY_FORCE: per slab node the spatial Gaussian exp(-d^2 / R^2) about
    // the perturbation centre (host, glibc), -1 beyond three radii
    const double* s_psp;
    int pert;
    double pert_scale, pert_start, pert_dur;
    // ACCELERATION_ACTIVE: s_om holds rpm 2 pi / 60 unscaled and the device applies the
    // scaling factor of the iteration's physical time
    int accel;
    double acc_t0, acc_t1, acc_s0, acc_s1, om_scaling;
    double th_kn, th_kf, th_kd, th_nu;
    int th_off_design, stall;
    const double* axi_gf;              // axisymmetric: |S_K(i,j,0)| * wedge angle, (j*NI + i)
    // side outputs
    double *finv, *fvisc, *dev;
    // time step
    double* dt;                        // N
    double* dtmin;                     // [1]
    // convergence / control on device
    int* stop;                         // non-zero: skip all work
    long* iter;                        // current iteration counter (1-based during an iteration)
    ConvDev conv;
    // state buffers by role: 0 = A holds U^(0) at iteration start and receives the last
    // stage in place; stages alternate between B (1) and C (2) in between
    // laminar viscous terms (VISCOSITY_ACTIVE) and no-slip walls
    int visc, mu_const, wall_moving;   // wall_moving: a no-slip wall with a nonzero velocity
    double mu_c, mu_ref, mu_tref, mu_s, cp_v, pr;
    const int* nsw;                    // N: node on a no_slip_wall patch (null when there is none)
    const double* wallv;               // [3][N] wall velocity of those nodes
    const double* gw;                  // [6][N] face interpolation weights W E N S B T
    double* grad;                      // [ux, uy, uz, T][x, y, z][N] stage-input gradients
    double* vf;                        // [I, J, K][mom x, y, z, energy][N] viscous flux * |S| of
                                       // the face before each node (index i, j or k)

    double* U[3];
    int nstages;
    double alpha[4];
    // stage-kernel tiling: k planes per block and blocks along i, j, k
    int kc, sgx, sgy, sgz;
    double* rper;                      // [2][5][NI*NJ] residuals of the periodic planes
    double* partial;                   // [stage blocks + periodic blocks][5] norm partial sums
    int nparts_stage, nparts_per;
    double* partial_min;               // [nblk_dt] per-block minima of the time step (global dt, turbo time)
    // turbomachinery history (TURBO_OUTPUT): the inlet and outlet boundary of the turbo
    // direction, turbo_nf faces each (device node index, boundary surface vector)
    int turbo_on, turbo_nf, turbo_nblk;  // nblk: partial-sum blocks per boundary (256 faces each)
    double* turbo_part;                // [2][nblk][5] partial sums
    const int* turbo_node;             // [2][nf]
    const double* turbo_S;             // [2][nf][3]
    double turbo_mscale;               // 2 pi / wedge angle, 360 / periodicity angle, or 1
    double* turbo_hist;                // [hist_len][5]: time after the iteration, mass flow, PRtt, TRtt, ETAtt
    double* turbo_time;                // physical time accumulated over the iterations
    double* t_now;                     // physical time at the start of the current iteration
    int nblk_dt;
};

// One block of a launch: the case it works on and its block coordinates in that case.
struct Work {
    int cs, x, y, z;
};

// Kernel argument of every launch: all cases of a batch advance in the same launch.
struct Launch {
    const CaseDev* cases;
    const Work* work;
    int in, out;           // state-buffer roles read and written
    int stage;             // RK stage index
    int last;              // last stage: norms and side outputs
    int rmode;             // write the residual instead of the update (residual entry point)
    int pre;               // first stage of an iteration: moving-wall viscous cases read the
                           // preprocessed input (U[2], wall momentum = rho * wall velocity)
};

// The state a stage reads: the iteration's first stage of a viscous case with a moving
// no-slip wall reads the copy whose wall momentum was reset (solver.cpp:1365, 3189-3204).
__device__ __forceinline__ const double* stage_input(const CaseDev& c, const Launch& La) {
    return (La.pre && c.wall_moving) ? c.U[2] : c.U[La.in];
}

// RK stage s of ns reads role stage_in(s) and writes role stage_out(s, ns):
// A -> B -> C -> B -> A (4 stages), A -> B -> C -> A (3 stages).
__host__ __device__ inline int stage_in(int s) { return s == 0 ? 0 : (s % 2 == 1 ? 1 : 2); }
__host__ __device__ inline int stage_out(int s, int ns) { return s == ns - 1 ? 0 : (s % 2 == 0 ? 1 : 2); }

// Global-memory access by a uniform base pointer and a 32-bit element index.  The device
// arrays live in global memory, but pointers read from the case table are generic ones,
// which the compiler would access with flat instructions (slower, and counted with LDS
// operations so that every LDS wait also waits for them).  The cast makes them global
// loads and stores with the base in scalar registers and a 32-bit byte offset per lane;
// the library keeps every array below 4 GiB.
#define TBFM_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ T gld(const T* base, unsigned idx) {
    return *(const TBFM_GLOBAL T*)((const TBFM_GLOBAL char*)base + idx * (unsigned)sizeof(T));
}
// A load with the non-temporal hint, for data each kernel reads once (face records, dual
// grid points): it is marked for early eviction, which keeps the rest of the working
// set (state buffers, source buffer, blade-row data) in the L2 and the Infinity Cache
// between kernels.
template <typename T>
__device__ __forceinline__ T gldnt(const T* base, unsigned idx) {
    return __builtin_nontemporal_load((const TBFM_GLOBAL T*)((const TBFM_GLOBAL char*)base + idx * (unsigned)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ void gst(T* base, unsigned idx, T v) {
    *(TBFM_GLOBAL T*)((TBFM_GLOBAL char*)base + idx * (unsigned)sizeof(T)) = v;
}

// The 32-byte record of face `rec` (unit normal and area) as two 16-byte loads, with the
// non-temporal hint (the records are read once per stage).
typedef double dpair __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void gld_face(const double* base, unsigned rec, double out[4]) {
    const TBFM_GLOBAL dpair* p = (const TBFM_GLOBAL dpair*)((const TBFM_GLOBAL char*)base + rec * 32u);
    const dpair a = __builtin_nontemporal_load(p), b = __builtin_nontemporal_load(p + 1);
    out[0] = a.x; out[1] = a.y; out[2] = b.x; out[3] = b.y;
}

typedef float fquad __attribute__((ext_vector_type(4)));
template <typename RT>
__device__ __forceinline__ void gld_face32(const float* base, unsigned rec, RT out[4]) {
    const fquad a = __builtin_nontemporal_load((const TBFM_GLOBAL fquad*)((const TBFM_GLOBAL char*)base + rec * 16u));
    out[0] = a.x; out[1] = a.y; out[2] = a.z; out[3] = a.w;
}
__device__ __forceinline__ double dmaxd(double a, double b) { return (a < b) ? b : a; }
// std::max of either precision (the first argument unless it is less than the second)
template <typename T>
__device__ __forceinline__ T smax(T a, T b) { return (a < b) ? b : a; }
__device__ __forceinline__ double dmind(double a, double b) { return (b < a) ? b : a; }

}  // namespace dev
}  // namespace tbfm
