/* CPU restatement of CTurboBFM's explicit residual path -- TEST INFRASTRUCTURE.
 *
 * This is the oracle: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the test oracle.  It is never part of the
 * product path.  Plain C11, compiled with -ffp-contract=off; OpenMP over grid lines
 * only, with the per-node operation order of the reference kept unchanged.
 *
 * Layout follows the reference: scalar node arrays are k-fastest,
 * idx = (i*NJ + j)*NK + k (include/types.hpp:411,428); conserved state is five such
 * arrays (rho, rhoU, rhoV, rhoW, rhoE), include/types.hpp:763-1018.
 */
#ifndef TBFM_ORACLE_H
#define TBFM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { TBO_TOPO_2D = 0, TBO_TOPO_3D = 1, TBO_TOPO_AXI = 2, TBO_TOPO_1D = 3 };
enum { TBO_JST = 0, TBO_ROE = 1 };
enum { TBO_LIM_NONE = 0, TBO_LIM_VAN_ALBADA = 1, TBO_LIM_VAN_LEER = 2, TBO_LIM_MIN_MOD = 3 };
enum { TBO_DT_LOCAL = 0, TBO_DT_GLOBAL = 1, TBO_DT_FIXED = 2 };
enum { TBO_BFM_NONE = 0, TBO_BFM_HALL = 1, TBO_BFM_LIFT_DRAG = 2, TBO_BFM_BLOCKAGE = 3, TBO_BFM_THOLLET = 4 };
enum { TBO_BC_INLET = 0, TBO_BC_OUTLET = 1, TBO_BC_RADIAL_EQ = 2, TBO_BC_WALL = 3,
       TBO_BC_PERIODIC = 4, TBO_BC_WEDGE = 5, TBO_BC_THROTTLE = 6, TBO_BC_MASS_FLOW = 7,
       TBO_BC_TRANSPARENT = 8, TBO_BC_NO_SLIP = 9 };

typedef struct {
    long ni, nj, nk;
    int topology, scheme, muscl, limiter;
    int n_stages;
    double rk[4];
    int dt_method;
    double cfl, fixed_dt;
    double gamma, r_gas;
    double init_mach, init_temperature, init_pressure, init_dir[3];
    double ramp_iterations;          /* OUTLET_PRESSURE_RAMP_ITERATIONS (0 = none) */
    int bfm_active, bfm_model, blockage_active;
    long le_index, te_index;         /* LEADING/TRAILING_EDGE_INDEX (lift-drag) */
    double omega_scaling;            /* ROTATIONAL_SPEED_SCALING_FACTOR */
    int inlet_cylindrical;           /* INLET_REFERENCE_FRAME = cylindrical */
    int turbo_dir;                   /* 0: i, 1: -i, 2: j, 3: -j */
    long residual_drop;              /* RESIDUALS_DROP_CONVERGENCE */
    int steady;
    int fluid_real;                  /* FLUID_MODEL = real: look-up-table fluid (tbo_set_lut) */
    /* Hall-Thollet (include/source_bfm_thollet.hpp:47-50, config.hpp:295-313) */
    double kn, kf, kd;               /* HALL_THOLLET_COEFFICIENT_KN / KF / KD */
    int thollet_off_design;          /* HALL_THOLLET_OFF_DESIGN_ACTIVE */
    int stall_active;                /* STALLED_BFM_ACTIVE (all models' force direction) */
    double nu;                       /* FLUID_KINEMATIC_VISCOSITY */
    /* mass-flow outlet (config.hpp:291-293) */
    double mass_flow_relaxation;     /* MASS_FLOW_RELAXATION (0.05) */
    double mass_flow_max_change;     /* MASS_FLOW_MAX_PRESSURE_CHANGE (0.02) */
    /* Greitzer plenum (greitzer_model.cpp, config.hpp:315-353) */
    int greitzer;                    /* ENABLE_GREITZER_MODELING (fixed time step) */
    double greitzer_volume;          /* GREITZER_PLENUM_VOLUME */
    double greitzer_ambient;         /* GREITZER_AMBIENT_PRESSURE (101325) */
    /* laminar viscous terms (VISCOSITY_ACTIVE, config.hpp:341-377, fluid_ideal.cpp:209-235) */
    int viscous;                     /* VISCOSITY_ACTIVE */
    int mu_constant_model;           /* FLUID_VISCOSITY_MODEL constant (else Sutherland) */
    double mu_constant;              /* FLUID_MU_CONSTANT */
    double mu_ref, mu_t_ref, mu_s;   /* FLUID_SUTHERLAND_MU_REF / _T_REF / _S */
    double cp_visc;                  /* FLUID_CP (thermal conductivity only) */
    double prandtl;                  /* FLUID_PRANDTL_NUMBER */
    /* time-dependent rotational speed (ACCELERATION_ACTIVE, fixed time step; config.cpp:695-718) */
    int accel_active;
    double acc_t0, acc_t1;           /* ACCELERATION_INITIAL_TIME / _FINAL_TIME */
    double acc_s0, acc_s1;           /* ROTATIONAL_SPEED_SCALING_FACTOR_INITIAL / _FINAL */
    /* Gaussian body-force perturbation (PERTURBATION_BODY_FORCE, source_bfm_base.cpp:215-243) */
    int pert_active;
    double pert_center[3];           /* PERTURBATION_CENTER */
    double pert_radius;              /* PERTURBATION_RADIAL_EXTENSION */
    double pert_scale;               /* PERTURBATION_SCALING_FACTOR */
    double pert_start, pert_duration; /* PERTURBATION_TIME_START / _TIME_DURATION */
} tbo_params;

/* A fluid look-up table (include/fluid_real.hpp:11-60 Table2D): nx*ny values, row-major
 * data[i*ny + j] over a uniform grid x in [xmin, xmax], y in [ymin, ymax]. */
typedef struct {
    long nx, ny;
    double xmin, xmax, ymin, ymax;
    const double* data;
} tbo_table;

/* The tables of a FluidReal (src/fluid_real.cpp:12-209), in this order:
 * RHO_E (x = rho, y = e): p, T, a, s, mu, kappa, dp/drho|e, dp/de|rho;
 * P_T (x = p, y = T): rho, e, a, s;  P_S (x = p, y = s): rho, T, e. */
enum { TBO_LUT_P, TBO_LUT_T, TBO_LUT_A, TBO_LUT_S, TBO_LUT_MU, TBO_LUT_KAPPA, TBO_LUT_DPDRHO,
       TBO_LUT_DPDE, TBO_LUT_RHO_PT, TBO_LUT_E_PT, TBO_LUT_A_PT, TBO_LUT_S_PT, TBO_LUT_RHO_PS,
       TBO_LUT_T_PS, TBO_LUT_E_PS, TBO_LUT_COUNT };
typedef struct {
    tbo_table t[TBO_LUT_COUNT];
    double gamma_nominal;            /* NOMINAL_GAMMA */
} tbo_lut;

typedef struct {
    int type;
    int nvalues;
    double values[8];
    long i0, i1, j0, j1, k0, k1;
} tbo_patch;

/* Static per-node input fields (grid CSV columns); any may be NULL if absent. */
typedef struct {
    const double *blockage, *normal_axial, *normal_radial, *normal_tangential;
    const double *rpm, *blade_present, *number_blades, *metal_angle;
    const double *streamwise_length, *lift_drag_beta0, *lift_drag_kp;
    const double *deviation_angle_pivot, *deviation_angle_stall;
} tbo_fields;

typedef struct tbo_case tbo_case;

tbo_case* tbo_create(const tbo_params* p, const double* x, const double* y, const double* z,
                     const tbo_fields* f, const tbo_patch* patches, int npatches,
                     char* err, int errlen);
void tbo_destroy(tbo_case* c);
/* Installs the look-up-table fluid of a FLUID_MODEL = real case (copies the tables);
 * call before tbo_init_state. */
void tbo_set_lut(tbo_case* c, const tbo_lut* lut);
/* FluidReal point functions of an installed table, for per-function pinning:
 * which 0 p(rho,e) 1 T(rho,e) 2 a(rho,e) 3 s(rho,e) 4 dp/drho|e 5 dp/de|rho 6 e(p,rho)
 * 7 a(p,rho) 8 rho(p,T) 9 s(p,T) 10 rho(p,s) 11 T(p,s) 12 e(p,s) 13 h(p,s)
 * 14 e(p,T) 15 a(p,T) 16 p(rho,T); for (x, y) in that order. */
double tbo_lut_eval(const tbo_case* c, int which, double x, double y);
long tbo_num_nodes(const tbo_case* c);

/* state: 5*N doubles, variable-major (rho[N], rhoU[N], ...), k-fastest */
void tbo_init_state(tbo_case* c);          /* initializeSolutionFromScratch + radial profiles */
void tbo_set_state(tbo_case* c, const double* U, int reset_profiles);
void tbo_get_state(const tbo_case* c, double* U);
long tbo_get_profile(const tbo_case* c, int q, double* out);

/* Runs n explicit iterations of Solver::solve (solver.cpp:1337-1424).
 * logres: n*5 log10 residual norms; turbo: n*4 {massflow, PRtt, TRtt, eta} at the
 * top of each iteration (may be NULL).  Returns the iteration at which the
 * convergence test passed (0 if it never did), or -1-v if residual v diverged. */
long tbo_run(tbo_case* c, long n, double* logres, double* turbo, int nthreads);

/* One computeResiduals(U) at iteration it with the current time-step array; R is
 * 5*N.  Refreshes the radial-equilibrium profiles first when update_profiles. */
void tbo_residual(tbo_case* c, const double* U, long it, int update_profiles, double* R);
/* The source terms alone (solver.cpp:2790-2897): R = minus the terms selected by which
 * (bit 1 geometric, 2 Gong, 4 body force). */
void tbo_source_residual(tbo_case* c, const double* U, int which, double* R);
/* Test hooks: overwrite outlet profile q (prof_n values); the boundary flux of patch
 * pidx at face {i, j, k} as the residual computes it. */
long tbo_set_profile(tbo_case* c, int q, const double* in);
void tbo_patch_flux(const tbo_case* c, int pidx, const double* Ui, const double* S, const long* face, long it,
                    double* out);
/* Physical time of the next iteration (Solver::_currentTime, solver.cpp:41, 1389): the
 * sum of the previous iterations' minimum time steps. */
void tbo_set_time(tbo_case* c, double t);
double tbo_get_time(const tbo_case* c);
/* The Greitzer plenum alone (greitzer_model.cpp:16-56): state (p0, mout0, T0), then one
 * computePlenumPressure per mass flow. */
void tbo_greitzer_steps(tbo_case* c, double p0, double mout0, double T0, const double* mflow, int n,
                        double* pout, double* mout);
void tbo_timestep(tbo_case* c, const double* U, double* dt);

/* Geometry: which 0..2 surfaces I/J/K, 3..5 centres I/J/K, 6 volumes, 7 dual nodes,
 * 8 blockage gradient.  Same shapes as the reference Mesh arrays. */
int tbo_geometry(const tbo_case* c, int which, double* out);

/* Velocity and temperature gradients of the last residual evaluation of a viscous case:
 * 12*N, [ux, uy, uz, T][x, y, z][N] (Solver::computeSolutionGradient, solver.cpp:3033-3058). */
void tbo_gradients(const tbo_case* c, double* out);
/* Solver::computeViscousFlux of a conserved state U (5), gradients g (ux, uy, uz, T as
 * xyz triples, 12) and face vector S: the flux per unit area (5). */
void tbo_viscous_flux(const tbo_case* c, const double* U, const double* g, const double* S, double* out);

/* Side outputs of the last source evaluation: finv, fvisc (3*N xyz), deviation (N). */
void tbo_side_outputs(const tbo_case* c, double* finv, double* fvisc, double* dev, double* dt);

/* Turbo performance of a state (solver.cpp:1633-1781): out[0..5] boundary mass flows,
 * out[6] mass flow, out[7] PRtt, out[8] TRtt, out[9] efficiency. */
void tbo_turbo(tbo_case* c, const double* U, double* out);

/* Point functions, exposed for per-function pinning against the reference build. */
void tbo_face_flux(const tbo_case* c, const double* Ull, const double* Ul, const double* Ur,
                   const double* Urr, const double* S, double* out);
void tbo_boundary_flux(const tbo_case* c, int kind, const double* values, const double* Uint,
                       const double* S, const long* face, long iter, const double* profile,
                       double* out);
void tbo_bfm_source(tbo_case* c, long i, long j, long k, const double* primitive, double* out,
                    double* finv, double* fvisc, double* dev);
void tbo_fluid_state(const tbo_case* c, const double* primitive, double* out);
void tbo_radial_equilibrium(const double* rho, const double* ut, const double* r, int n,
                            double hub, double* p);

#ifdef __cplusplus
}
#endif
#endif
