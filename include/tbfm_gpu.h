/* C-ABI of the MI355X explicit residual path (JST / Roe-MUSCL fluxes, Hall / lift-drag /
 * Hall-Thollet body forces with the stalled-blade model, local / global / fixed time step,
 * RK41A / RK31B stage update).
 *
 * The reference has no FFI; its operator seams are per-element C++ virtuals
 * (AdvectionBase::computeFlux, include/advection_base.hpp:21-25;
 * BoundaryBase::computeBoundaryFlux, include/boundary_base.hpp:26-32;
 * SourceBFMBase::computeTotalSource, include/source_bfm_base.hpp:20-31) called from
 * Solver::computeResiduals (src/solver.cpp:1783-1833) inside Solver::solve
 * (src/solver.cpp:1298-1424).  A GPU cannot sit behind per-face virtual calls, so the
 * boundary is whole-array at Solver level: a case is created once from the same inputs
 * the reference reads (grid CSV columns, input.ini keys, boundaries CSV rows) and is
 * then advanced N explicit iterations on the device.
 *
 * Conventions
 *   - Host arrays use the reference's layout: node index (i*NJ + j)*NK + k (k fastest,
 *     include/types.hpp:411,428); a conserved state is 5 such arrays back to back
 *     (rho, rhoU, rhoV, rhoW, rhoE).  The library transposes to its device layout.
 *   - Host pointers are only read or written during the call; none is retained.
 *   - Every entry point returns TBFM_OK (0) or a negative code and never throws; the
 *     message of the last failure is tbfm_gpu_last_error(case) (or NULL case for
 *     creation failures).
 *   - One host thread may use a case (or batch) at a time; each case and batch owns its
 *     HIP stream, and different cases and batches may be driven from different host
 *     threads concurrently (allocation, stream capture and synchronous copies are
 *     serialised inside the library; graph replays overlap).
 *   - Fluid: FLUID_MODEL = ideal (FLUID_GAMMA, FLUID_R_CONSTANT) or real, the look-up
 *     table of FLUID_TABLE_FILE (default fluid_table_<FLUID_NAME>.lut; a relative path
 *     is taken from the input.ini folder, or from the working folder for a case built
 *     from a descriptor) in the reference's .lut format (src/fluid_real.cpp:12-209).  The table fluid computes in FP64 (GPU_PRECISION = mixed is refused).
 */
#ifndef TBFM_GPU_H
#define TBFM_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define TBFM_OK 0
#define TBFM_ERR_ARGUMENT (-1)
#define TBFM_ERR_HIP (-2)
#define TBFM_ERR_OUT_OF_MEMORY (-3)
#define TBFM_ERR_DIVERGED (-4)
#define TBFM_ERR_UNSUPPORTED (-5)

typedef struct tbfm_case tbfm_case;
typedef struct tbfm_batch tbfm_batch;

/* Everything the reference reads from disk, passed as arrays.
 *   x, y, z, field_data[f]: N = ni*nj*nk doubles each, reference order (k fastest);
 *     field names are the grid-CSV column names (types.hpp:1084-1126), e.g.
 *     "blockage", "normalAxial", "rpm", "bladePresent", "numberBlades".
 *   config_keys/values: input.ini KEY = value pairs (config.cpp:15-44), including the
 *     patch lines NAME = type, v1, v2, ...
 *   patch_names / patch_ranges: the boundaries-CSV rows (solver.cpp:232-294), six
 *     longs per patch: I_MIN, I_MAX, J_MIN, J_MAX, K_MIN, K_MAX. */
typedef struct {
    long ni, nj, nk;
    const double* x;
    const double* y;
    const double* z;
    int nfields;
    const char* const* field_names;
    const double* const* field_data;
    int nconfig;
    const char* const* config_keys;
    const char* const* config_values;
    int npatches;
    const char* const* patch_names;
    const long* patch_ranges;
} tbfm_case_desc;

/* Replaces Solver::Solver(config, mesh) (src/solver.cpp:18-31) for the GPU path:
 * geometry (Mesh, src/mesh.cpp:11-40), boundary maps (solver.cpp:491-561), BFM set-up
 * (solver.cpp:701-749) and the initial solution (solver.cpp:777-888). */
int tbfm_gpu_case_create(const tbfm_case_desc* desc, tbfm_case** out);

/* Same, reading input.ini, GRID_FILE and BOUNDARY_CONDITIONS_FILEPATH from disk like
 * `turbobfm input.ini` (src/turbobfm.cpp:7-39). */
int tbfm_gpu_case_create_from_ini(const char* ini_path, tbfm_case** out);

int tbfm_gpu_case_destroy(tbfm_case* c);

/* Last error message of a case, or of the last failed creation when c is NULL. */
const char* tbfm_gpu_last_error(const tbfm_case* c);

void tbfm_gpu_dims(const tbfm_case* c, long* ni, long* nj, long* nk);

/* Conserved state (5*N doubles, reference layout).  set_state also re-derives the
 * radial-equilibrium outlet profiles from the new state, as the reference does after
 * initialisation (solver.cpp:818-834). */
int tbfm_gpu_set_state(tbfm_case* c, const double* U);
int tbfm_gpu_get_state(tbfm_case* c, double* U);

/* Replaces the explicit branch of Solver::solve (solver.cpp:1337-1424): runs up to
 * n_iter iterations (time step, radial profiles, RK stages with periodic closure,
 * residual norms, convergence test).  logres (may be NULL) receives n_iter*5 values of
 * log10(||R_v||_2 / N) (solver.cpp:1488-1508); rows after convergence are left
 * untouched.  *converged_at is the iteration (counted from the case's start) at which
 * the reference's convergence test (solver.cpp:2985-3015) passed, 0 if it has not.
 * Returns TBFM_ERR_DIVERGED when a norm is NaN/Inf, like the reference's
 * "Solver diverged" exception. */
int tbfm_gpu_advance(tbfm_case* c, long n_iter, double* logres, long* converged_at);

/* One Solver::computeResiduals (solver.cpp:1783-1833) of the given state at iteration
 * `it`, with the time step of that state; R is 5*N (reference layout).  When
 * update_profiles is non-zero the radial-equilibrium profiles are refreshed from U first
 * (solver.cpp:2688-2787).  Used for per-stage parity. */
int tbfm_gpu_compute_residual(tbfm_case* c, const double* U, long it, int update_profiles, double* R);

/* Solver::computeTimestepArray (solver.cpp:1531-1630) of the given state; dt is N. */
int tbfm_gpu_compute_timestep(tbfm_case* c, const double* U, double* dt);

/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * the time step of tbfm_gpu_compute_timestep for a state already in device memory,
 * computed without allocating.  dU: 5 planes of N doubles in the library's device
 * layout (node (k*NJ + j)*NI + i, i fastest); ddt: N doubles, same layout (the global
 * time step writes the domain minimum at every node).  Kernels are enqueued on `stream`
 * (a hipStream_t; NULL: the case's stream) and the call returns without waiting.  The
 * per-block minima of the global time step live in a file-scope device array of
 * TBFM_NOALLOC_MAX_BLOCKS + 1 doubles (512 KiB per GPU, held until the process ends), so
 * grids above TBFM_NOALLOC_MAX_BLOCKS * 256 nodes are refused; calls that overlap on one
 * GPU share that array and race. */
#define TBFM_NOALLOC_MAX_BLOCKS 65536
int tbfm_gpu_compute_timestep_noalloc(tbfm_case* c, const double* dU, double* ddt, void* stream);

/* Side outputs the reference keeps (solver.hpp:370-371): inviscid and viscous body
 * force (3*N, xyz per node), deviation angle (N) and time step (N); any may be NULL. */
int tbfm_gpu_get_side_outputs(tbfm_case* c, double* finv, double* fvisc, double* dev, double* dt);

/* Solver::updateMassFlows + updateTurboPerformance (solver.cpp:1633-1781) of the
 * current state: out[0..5] boundary mass flows (I/J/K start/end), out[6] mass flow,
 * out[7] total-pressure ratio, out[8] total-temperature ratio, out[9] efficiency. */
int tbfm_gpu_get_turbo(tbfm_case* c, double* out);

/* Turbomachinery performance of every iteration, recorded on the device when the case
 * was created with TURBO_OUTPUT = yes (Solver::updateMassFlows + updateTurboPerformance
 * at the start of each iteration, solver.cpp:1340-1341, 1633-1781; the physical time
 * accumulated as solver.cpp:1383-1385).  Rows of 5 values for iterations first ..
 * first+n-1 (counted from 1): time after the iteration [s], mass flow [kg/s], PRtt,
 * TRtt, ETAtt.  The boundary sums are tree reductions (the reference sums in face order:
 * the values agree to the last bits, and exactly at the six digits turbo.csv keeps). */
int tbfm_gpu_get_turbo_history(tbfm_case* c, long first, long n, double* out);

/* Static geometry as uploaded (reference layout), for parity checks of the mesh:
 * which 0..2 surfaces I/J/K, 3 volumes, 4 blockage gradient. */
int tbfm_gpu_get_geometry(tbfm_case* c, int which, double* out);

/* Timing of the fused residual+RK stage kernel: when enabled, every stage is bracketed
 * by HIP events on the case stream, one before the pre-stage kernel (boundary fluxes,
 * blade-row sources), one between it and the stage kernel, one after the stage kernel.
 * get returns the accumulated stage-kernel milliseconds and launch count since the last
 * reset, plus the algorithmic bytes of one stage-kernel launch (per-node contract times
 * nodes); get_prestage the pre-stage milliseconds over the same launches and the
 * body-force data bytes one pre-stage launch reads. */
int tbfm_gpu_profile(tbfm_case* c, int enable, int reset);
int tbfm_gpu_profile_get(tbfm_case* c, double* stage_ms, long* launches, double* bytes_per_launch);
int tbfm_gpu_profile_get_prestage(tbfm_case* c, double* prestage_ms, double* bytes_per_launch);

/* Measurement helper with no counterpart in the reference: runs `repeats` plain
 * device-to-device copies of nbytes (kernel k_copy_calibration) on the default stream
 * and returns the average milliseconds per copy.  rocprofv3's FETCH_SIZE/WRITE_SIZE of
 * this kernel, against its known nbytes read and written, scales the counters of the
 * residual kernels to HBM bytes. */
int tbfm_gpu_copy_calibration(long nbytes, int repeats, double* ms_per_copy);

/* Report of the stage-kernel autotuner (tile shape and k planes per block, timed on
 * the case's own state before its first iteration); empty before the first advance. */
const char* tbfm_gpu_tuning(const tbfm_case* c);

/* Synchronises the case stream. */
int tbfm_gpu_sync(tbfm_case* c);

/* ---------------------------------------------------------------- batches and serving
 * No counterpart in the reference, which runs one case per process (apps/turbobfm.cpp,
 * Solver::solve, src/solver.cpp:1298-1424).  A batch advances many cases (operating
 * points, grids of uneven size) together: every kernel of an iteration is one launch
 * over all members, through device work tables, and the iteration is replayed as one
 * captured HIP graph.  Each member keeps its own device memory, state, iteration
 * counter, residual history and convergence test, so a case computes exactly the same
 * numbers alone and in any mix.  Cases keep ownership: retire or destroy them freely
 * between advances; a member is advanced only through its batch. */

/* Free and total device memory of the current GPU (hipMemGetInfo): the budget a
 * serving loop admits cases against. */
int tbfm_gpu_device_memory(long* free_bytes, long* total_bytes);

/* Device memory a case holds (state, geometry, boundary and reduction buffers): the
 * admission footprint of a serving loop. */
int tbfm_gpu_case_device_bytes(const tbfm_case* c, long* bytes);

/* Replaces the convergence rule of Solver::checkConvergence (src/solver.cpp:2985-3015)
 * for this case: drop = orders of magnitude below the baseline of iterations 1-5;
 * density_only = 1 tests the density residual alone (the serving criterion). */
int tbfm_gpu_case_set_convergence(tbfm_case* c, int density_only, double drop);

/* Iterations executed so far, convergence iteration (0 if none), whether the case has
 * stopped (converged or diverged), and the log residuals of its last iteration;
 * TBFM_ERR_DIVERGED after a NaN/Inf residual (solver.cpp:1494-1496). */
int tbfm_gpu_case_status(tbfm_case* c, long* iterations, long* converged_at, int* stopped,
                         double* last_logres);

int tbfm_gpu_batch_create(tbfm_batch** out);
/* Admits a case (autotunes its tiling on first use); members must share the RK scheme. */
int tbfm_gpu_batch_admit(tbfm_batch* b, tbfm_case* c);
int tbfm_gpu_batch_retire(tbfm_batch* b, tbfm_case* c);
int tbfm_gpu_batch_size(const tbfm_batch* b);
/* Advances every member n iterations (stopped members are skipped on the device) and
 * synchronises. */
int tbfm_gpu_batch_advance(tbfm_batch* b, long n_iter);
/* Retires every member (the cases stay valid) and frees the batch. */
int tbfm_gpu_batch_destroy(tbfm_batch* b);
const char* tbfm_gpu_batch_last_error(const tbfm_batch* b);

/* ---- on-disk formats */

/* Output::writeSolution (src/output.cpp:367-447) of the case's current state and side
 * outputs: "NI=", "NJ=", "NK=" lines, the header "x,y,z" plus the field names of
 * OUTPUT_FIELDS_TYPE (primary; secondary, the default; turbo_bfm) in name order, one row
 * per node with k fastest, 16 significant digits.  path is the file to write (the
 * reference writes Volume_CSV/<SOLUTION_NAME>[_<iteration>].csv, see
 * tbfm_volume_file_name). */
int tbfm_gpu_write_solution(tbfm_case* c, const char* path);

/* The same file for a state given by the caller (U: 5*N, finv / fvisc: 3*N xyz
 * triples, dev: N, reference layout; the side outputs may be NULL for zeros), with the
 * grid, fluid and output keys read from input.ini.  Host only: no device is used. */
int tbfm_write_solution_from_ini(const char* ini_path, const double* U, const double* finv, const double* fvisc,
                                 const double* dev, const char* path);

/* Output::getOutputFilename (src/output.cpp:352-364): SOLUTION_NAME, with
 * "_<6-digit iteration>" when SAVE_UNSTEADY (default yes); written into name (size
 * bytes, NUL-terminated).  Host only. */
int tbfm_volume_file_name(const char* ini_path, long iteration, char* name, long size);

/* Initial state the case created from input.ini starts from (replaces
 * Solver::initializeSolutionArrays' state part, solver.cpp:777-1295): the uniform
 * INIT_* state (INIT_DIRECTION = 0,0,0: the grid's i-direction per node), or with
 * RESTART_SOLUTION = yes the restart file RESTART_SOLUTION_FILEPATH: same dimensions ->
 * standard restart; RESTART_TYPE = axisymmetric with equal NI, NJ -> the k = 0 plane
 * rotated to every node; nearest_neighbor[_axisymmetric] or other dimensions -> the
 * nearest file node (in x, y, z, or in x, r with the velocity rotated).  U is 5*N;
 * finv / fvisc (3*N, may be NULL) receive the restart's body forces (zeros otherwise).
 * Host only. */
int tbfm_initial_state_from_ini(const char* ini_path, double* U, double* finv, double* fvisc);

#ifdef __cplusplus
}
#endif
#endif
