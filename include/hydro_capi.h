/*
 * hydro_capi.h -- C ABI of the MI355X finite-volume update path.
 *
 * This is the drop-in boundary for the update path of KennethYeah/HydrodynamicsGPU:
 * the reference's C++ classes (Solver, Roe, HLL, EulerRoe, EulerHLL, EulerHLLC,
 * SRHDRoe, MaxwellRoe, Integrator/ForwardEuler/RungeKutta<...>) call these entry
 * points instead of binding OpenCL kernels by index.  Only plain C types cross
 * the boundary: ints, doubles, pointers and sizes.
 *
 * Conventions shared by every entry point
 *   - Status: 0 = ok, negative = error (HYDRO_E_*).  hydro_last_error() returns
 *     a thread-local message for the last failing call on this thread.
 *   - Host arrays are borrowed for the duration of the call only.
 *   - State arrays at the boundary use the reference layout: array-of-structs
 *     state[s + N*cell], cell = x + SX*(y + SY*z), ghost cells included
 *     (res/include/HydroGPU/Shared/Common.h:59-60).  On the device the library
 *     stores structure-of-arrays planes; the conversion happens in the copy.
 *   - One HIP stream per hydro_sim.  A handle is not re-entrant; distinct
 *     handles may be driven from different threads.
 */
#ifndef HYDRO_CAPI_H
#define HYDRO_CAPI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enums --------------------------------------------------------------- */

/* status codes */
#define HYDRO_OK 0
#define HYDRO_E_INVALID (-1)   /* bad argument / unsupported combination */
#define HYDRO_E_HIP (-2)       /* HIP runtime error */
#define HYDRO_E_NUMERIC (-3)   /* non-finite dt, negative density/pressure, ... */
#define HYDRO_E_UNSUPPORTED (-4)

/* equations (src/Equation/{Euler,SRHD,Maxwell}.cpp) */
enum { HYDRO_EQN_EULER = 0, HYDRO_EQN_SRHD = 1, HYDRO_EQN_MAXWELL = 2 };

/* flux solvers (src/Solver/{Roe,HLL}.cpp, res/EulerHLLC.cl; EulerBurgers:
 * src/Solver/EulerBurgers.cpp, res/EulerBurgers.cl -- Euler only) */
enum { HYDRO_SOLVER_ROE = 0, HYDRO_SOLVER_HLL = 1, HYDRO_SOLVER_HLLC = 2, HYDRO_SOLVER_BURGERS = 3 };

/* slope limiters, in the order of res/SlopeLimiter.cl:5-68 */
enum {
  HYDRO_LIMITER_DONORCELL = 0,
  HYDRO_LIMITER_LAXWENDROFF,
  HYDRO_LIMITER_BEAMWARMING,
  HYDRO_LIMITER_FROMM,
  HYDRO_LIMITER_CHARM,
  HYDRO_LIMITER_HCUS,
  HYDRO_LIMITER_HQUICK,
  HYDRO_LIMITER_KOREN,
  HYDRO_LIMITER_MINMOD,
  HYDRO_LIMITER_OSHKER,
  HYDRO_LIMITER_OSPRE,
  HYDRO_LIMITER_SMART,
  HYDRO_LIMITER_SWEBY,
  HYDRO_LIMITER_UMIST,
  HYDRO_LIMITER_VANALBADA1,
  HYDRO_LIMITER_VANALBADA2,
  HYDRO_LIMITER_VANLEER,
  HYDRO_LIMITER_MONOTIZEDCENTRAL,
  HYDRO_LIMITER_SUPERBEE,
  HYDRO_LIMITER_BARTHJESPERSEN,
  HYDRO_NUM_LIMITERS
};

/* integrators: the registry of src/Solver/Solver.cpp:178-204 minus the
 * unfinished BackwardEulerConjugateGradient (src/Integrator/BackwardEulerConjugateGradient.cpp:80) */
enum {
  HYDRO_INTEGRATOR_FORWARDEULER = 0,
  HYDRO_INTEGRATOR_RK2,
  HYDRO_INTEGRATOR_RK2HEUN,
  HYDRO_INTEGRATOR_RK2RALSTON,
  HYDRO_INTEGRATOR_RK3,
  HYDRO_INTEGRATOR_RK4,
  HYDRO_INTEGRATOR_RK4_3_8THSRULE,
  HYDRO_INTEGRATOR_RK2TVD,
  HYDRO_INTEGRATOR_RK2NONTVD,
  HYDRO_INTEGRATOR_RK3TVD,
  HYDRO_INTEGRATOR_RK4TVD,
  HYDRO_INTEGRATOR_RK4NONTVD,
  HYDRO_NUM_INTEGRATORS
};

/* boundary methods (include/HydroGPU/Boundary/Boundary.h:8-14) */
enum {
  HYDRO_BOUNDARY_NONE = -1,
  HYDRO_BOUNDARY_PERIODIC = 0,
  HYDRO_BOUNDARY_MIRROR = 1,
  HYDRO_BOUNDARY_FREEFLOW = 2
};

/* arithmetic / storage precision of the update path */
enum {
  HYDRO_PRECISION_FP64 = 0,       /* FP64 state, FP64 arithmetic */
  HYDRO_PRECISION_FP32_MIXED = 1  /* FP32 state + flux/eigen arithmetic; FP64 dt
                                     reduction, totals and SRHD con2prim */
};

/* behaviour flags */
#define HYDRO_FLAG_REF_DX_6DIGITS 0x1u  /* kernel dx = 6-significant-digit rendering
                                           (include/HydroGPU/toNumericString.h:12-22,
                                           src/Solver/Solver.cpp:272-280) */
#define HYDRO_FLAG_GRAVITY_IDENTITY_PASS 0x2u /* replay the zero-potential self-gravity
                                           RK pass of EulerRoe/EulerHLL and the zero
                                           source pass of MaxwellRoe
                                           (src/Solver/EulerRoe.cpp:31-34,
                                           src/Solver/MaxwellRoe.cpp:38-46) */
#define HYDRO_FLAG_KAHAN 0x4u           /* FP32_MIXED: compensated (Kahan) state update */
#define HYDRO_FLAG_CLIP_TO_TEND 0x8u    /* clip the last step so t reaches t_end exactly */
#define HYDRO_FLAG_REFERENCE 0x3u       /* REF_DX_6DIGITS | GRAVITY_IDENTITY_PASS */

/* ---- simulation descriptor ---------------------------------------------- */

/* One grid.  The fields restate the reference's Lua config keys
 * (res/config.lua, src/HydroGPUApp.cpp:164-230) and its kernel `defs`. */
typedef struct hydro_sim_desc {
  int dim;               /* 1..3 ("last non-1 size", src/HydroGPUApp.cpp:213-222) */
  int size[3];           /* cells per axis INCLUDING 2+2 ghosts; unused axes = 1 */
  double xmin[3];        /* domain, ghosts inside (src/HydroGPUApp.cpp:224-226) */
  double xmax[3];
  int equation;          /* HYDRO_EQN_* */
  int solver;            /* HYDRO_SOLVER_* */
  int limiter;           /* HYDRO_LIMITER_*  (Roe only; HLL/HLLC have none) */
  int integrator;        /* HYDRO_INTEGRATOR_* */
  int precision;         /* HYDRO_PRECISION_* */
  int boundary[3][2];    /* [axis][min,max] HYDRO_BOUNDARY_* */
  double cfl;            /* stored as float by the reference (include/HydroGPU/HydroGPUApp.h:84) */
  int use_fixed_dt;      /* useFixedDT */
  double fixed_dt;       /* fixedDT */
  double t_end;          /* used with HYDRO_FLAG_CLIP_TO_TEND */
  /* defs (res/config.lua:70-93) */
  double gamma;          /* idealGas_heatCapacityRatio */
  int srhd_solve_prim_max_iter;
  double srhd_solve_prim_stop_epsilon;
  double srhd_solve_prim_vel_epsilon;
  double srhd_solve_prim_pmin_epsilon;
  double srhd_rho_min, srhd_rho_max, srhd_eint_max;
  double srhd_D_min, srhd_D_max, srhd_tau_min, srhd_tau_max;
  double maxwell_permittivity, maxwell_permeability, maxwell_conductivity;
  unsigned flags;        /* HYDRO_FLAG_* */
  /* self-gravity (Euler solvers: src/Solver/SelfGravitation.cpp,
   * res/SelfGravitation.cl; useGravity, gaussSeidelMaxIter and the def
   * selfGrav_gravitationalConstant of res/config.lua:65-74) */
  int use_gravity;                         /* relax the potential from the density */
  int gauss_seidel_max_iter;               /* relaxation sweeps per potential solve */
  double self_grav_gravitational_constant; /* G */
} hydro_sim_desc;

typedef struct hydro_ctx hydro_ctx;
typedef struct hydro_sim hydro_sim;
typedef struct hydro_batch hydro_batch;

/* Fill *d with the reference defaults (res/config.lua: Superbee, ForwardEuler,
 * FREEFLOW everywhere, cfl 0.5, gamma 1.4, srhd_* defs). */
void hydro_sim_desc_default(hydro_sim_desc* d);

/* Last error message of this thread ("" if none). */
const char* hydro_last_error(void);

/* Library version string. */
const char* hydro_version(void);

/* ---- context ------------------------------------------------------------- */

/* Replaces CLCommon's device/queue pick (src/HydroGPUApp.cpp:250-279). */
int hydro_ctx_create(int device, hydro_ctx** out);
void hydro_ctx_destroy(hydro_ctx* ctx);

/* ---- one simulation ------------------------------------------------------ */

/* Replaces Solver::init (src/Solver/Solver.cpp:48-205): validates the
 * descriptor, allocates every device buffer once. */
int hydro_sim_create(hydro_ctx* ctx, const hydro_sim_desc* d, hydro_sim** out);
void hydro_sim_destroy(hydro_sim* sim);

/* Number of states N per cell, and number of cells (ghosts included). */
int hydro_sim_num_states(const hydro_sim* sim);
size_t hydro_sim_volume(const hydro_sim* sim);

/* Replaces Solver::Converter::toGPU / Solver::resetState's upload
 * (src/Solver/Solver.cpp:333-337, 350-392).  `aos` holds n = N*volume doubles
 * in the reference layout.  For SRHD the values are the Newtonian
 * (rho, rho v, E) the reference's initState writes, and the call then runs
 * initVariables (src/Solver/SRHDRoe.cpp:40-45). */
int hydro_sim_set_state(hydro_sim* sim, const double* aos, size_t n);

/* Replaces Solver::Converter::fromGPU (src/Solver/Solver.cpp:339-343). */
int hydro_sim_get_state(hydro_sim* sim, double* aos, size_t n);

/* The self-gravity half of the Euler solvers' resetState
 * (SelfGravitationBehavior::Converter::toGPU -> SelfGravitation::resetState,
 * src/Solver/SelfGravitation.cpp:35-97), after hydro_sim_set_state: the
 * potential phi0 = `pot` (one value per cell, NULL: zero), or the guess
 * -rho when use_gravity is set, in which case 100 x gauss_seidel_max_iter
 * relaxation sweeps of the Poisson equation follow; then every cell's total
 * energy gains phi0 (the initial potential, as the reference adds it).  With
 * a non-zero potential or use_gravity, each update adds the reference's
 * gravity pass (SelfGravitation::applyPotential, :99-124: per stage
 * gauss_seidel_max_iter sweeps when use_gravity, then the gravity source
 * res/SelfGravitation.cl:65-102), and the Roe, HLL and EulerBurgers kernels
 * read the potential (HLLC's do not, as in res/EulerHLLC.cl).  The
 * reference relaxes in place in whatever order work-items run; here each
 * sweep reads the previous sweep everywhere (Jacobi order), the
 * order-independent form of the same update.  Euler solvers only; not for
 * 3D HLL, batches or the allocation-free entry point. */
int hydro_sim_set_potential(hydro_sim* sim, const double* pot, size_t n);
/* The potential (one value per cell, ghost cells included). */
int hydro_sim_get_potential(hydro_sim* sim, double* pot, size_t n);
/* Solid cells: the solidBuffer upload of SelfGravitation::resetState
 * (src/Solver/SelfGravitation.cpp:56-75; one flag per cell, ghost cells
 * included, nonzero = solid; NULL or all zero clears them).  Read by the
 * solvers the reference builds with SOLID: EulerRoe (res/EulerRoe.cl:48-65
 * mirror in the eigenbasis, res/Roe.cl:38-44 CFL, 110-124 and 235-290 mirror
 * and limiter ratio, res/CalcFluxDeriv.cl:28-30 no derivative in solid cells)
 * and EulerBurgers (res/EulerBurgers.cl:42-47, 119-127, 209-236, 307-309,
 * 372-400, 425-447).  EulerHLL/HLLC accept the mask and ignore it, as their
 * kernels do.  Euler solvers only; the mask stays until replaced. */
int hydro_sim_set_solid(hydro_sim* sim, const unsigned char* solid, size_t n);
/* The solid flags (one per cell, 0 or 1; all 0 without a mask): the
 * solidBuffer read-back of SelfGravitationBehavior::Converter::fromGPU
 * (include/HydroGPU/Solver/SelfGravitationBehavior.h:83-88). */
int hydro_sim_get_solid(hydro_sim* sim, unsigned char* solid, size_t n);

/* SRHD primitives (rho, v, eInt) in the reference layout (src/Solver/SRHDRoe.cpp:9-13).
 * They are stored in the state's precision: FP64, or FP32 for the FP32-mixed
 * form (recovered in FP64 by con2prim, then rounded for the FP32 fluxes and
 * speeds that read them), widened to double here. */
int hydro_sim_get_primitives(hydro_sim* sim, double* aos, size_t n);

/* The stored form of the state: for compensated FP32 storage
 * (HYDRO_FLAG_KAHAN) the high words, widened to double, and the low words
 * (`lo`, may be NULL), whose sum hydro_sim_get_state returns; for the other
 * forms the state itself, and zeros in `lo`.  No counterpart in the
 * reference; checkpoints keep this form so a resumed run repeats the
 * original bit for bit. */
int hydro_sim_get_storage(hydro_sim* sim, double* hi, float* lo, size_t n);

/* Installs a state read back with hydro_sim_get_state or
 * hydro_sim_get_storage as it is (no Newtonian-to-SR conversion, unlike
 * hydro_sim_set_state): `lo` are the low words of compensated FP32 storage
 * (NULL: split as hydro_sim_set_state does; must be NULL for the other
 * forms), `prims` the SRHD primitives from hydro_sim_get_primitives (must
 * be NULL for the other equations; rounded to FP32 for the FP32-mixed
 * form, which gives back exactly what hydro_sim_get_primitives read).  Sets the simulated time and step count.
 * No counterpart in the reference, which has no checkpoints; it resumes a
 * run saved by the checkpoint writers (hydrodynamicsgpu_amd/io.py,
 * HydroGPU::Solver::Solver::saveCheckpoint). */
int hydro_sim_restore(hydro_sim* sim, const double* aos, const float* lo, const double* prims, size_t n, double t,
                      long long steps);

/* Replaces Solver::boundary (src/Solver/Solver.cpp:446-464; SRHD adds the
 * primitive buffer, src/Solver/SRHDRoe.cpp:79-96). */
int hydro_sim_boundary(hydro_sim* sim);

/* Replaces Roe::calcTimestep / HLL::initStep + HLL::calcTimestep
 * (src/Solver/Roe.cpp:88-92, src/Solver/HLL.cpp:225-232): freezes the
 * start-of-step eigen data and returns cfl * min cell dt (FP64). */
int hydro_sim_calc_timestep(hydro_sim* sim, double* dt_out);

/* No counterpart in the reference.  The CFL timestep of
 * hydro_sim_calc_timestep for a state the caller keeps in device memory,
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It calls
 * no allocating routine: its scratch (the CFL kernel's per-wave minima and
 * the clock of the dt tail) is a file-scope device array of
 * HYDRO_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.  Every
 * call on one GPU shares that array, so calls that overlap on one GPU race.
 *   d        descriptor as for hydro_sim_create (sizes include ghost cells)
 *   state    device pointer: N state planes, cell x + SX*(y + SY*z), plane
 *            stride `plane` elements (>= the cell count), double for FP64,
 *            float for FP32_MIXED (the high words of compensated storage);
 *            ghost cells filled (boundary applied), as the CFL reads them
 *   prims    SRHD: device pointer to the FP64 primitive planes (same stride);
 *            NULL for the other equations
 *   t        simulated time, for HYDRO_FLAG_CLIP_TO_TEND
 *   dt_out   device pointer, receives cfl * min cell dt (or the fixed dt)
 *   flags_out device pointer or NULL: health bit 0 (see hydro_sim_health) is
 *            OR-ed in
 *   stream   hipStream_t, or NULL for the null stream; nothing is waited on
 * Grids whose CFL pass would need more minima than the array holds are
 * rejected with HYDRO_E_INVALID (none of the supported kernel choices
 * needs more than 65536). */
#define HYDRO_NOALLOC_SCRATCH_BYTES (256u + 8u * 65536u)
int hydro_calc_timestep_noalloc(const hydro_sim_desc* d, const void* state, size_t plane, const double* prims,
                                double t, double* dt_out, unsigned* flags_out, void* stream);

/* Replaces Solver::step(dt) for the solver (src/Solver/Roe.cpp:94-99,
 * src/Solver/HLL.cpp:234-241) including the integrator stages
 * (include/HydroGPU/Integrator/RungeKutta.h:59-125,
 * src/Integrator/ForwardEuler.cpp:8-32) and the per-solver tail
 * (gravity/source pass, SRHD constrainState + updatePrimitives), and
 * advances the simulated time by dt.  The frozen eigen data (Roe basis, HLL
 * wave speeds) are those of the state the step starts from; the reference
 * reuses the basis of its last calcTimestep call, which is the same state
 * whenever each step follows a calcTimestep, as in Solver::update. */
int hydro_sim_step(hydro_sim* sim, double dt);

/* Replaces Solver::update (src/Solver/Solver.cpp:521-545):
 * boundary -> initStep -> calcTimestep (or fixedDT) -> step.  Writes the dt
 * used to *dt_out (may be NULL).  Synchronises the stream. */
int hydro_sim_update(hydro_sim* sim, double* dt_out);

/* Runs `nsteps` updates back to back without any host synchronisation: dt
 * stays on the device (no per-step readback).  Asynchronous on the sim's
 * stream; use hydro_sim_sync / hydro_sim_time to wait. */
int hydro_sim_run(hydro_sim* sim, int nsteps);

/* Waits for all work on the sim's stream. */
int hydro_sim_sync(hydro_sim* sim);

/* Simulated time, number of completed steps and the last dt (synchronises). */
int hydro_sim_time(hydro_sim* sim, double* t, long long* steps, double* last_dt);

/* FP64 domain-integrated conserved totals over the interior cells
 * (sum of state * cell volume); totals_out has N entries. */
int hydro_sim_totals(hydro_sim* sim, double* totals_out);

/* Non-zero per-step health flags accumulated since creation:
 * bit0 non-finite or non-positive dt, or a non-finite wave speed in the CFL;
 * bit2 con2prim not converged. */
int hydro_sim_health(hydro_sim* sim, unsigned* flags_out);

/* No counterpart in the reference (its kernels' work sizes are fixed by the
 * OpenCL NDRange).  The work decomposition of the fused stage kernels:
 * params_out[4] = rows per wave (2D) or rows / planes per chunk (3D), waves
 * per workgroup (2D), the first-order 3D form (0 y-march, 1 z-march one cell
 * per lane, 2 z-march two cells per lane, 3 z-march one cell per lane in
 * 8-wave workgroups) and that form's waves per workgroup. */
int hydro_sim_decomposition(hydro_sim* sim, int* params_out);

/* No counterpart in the reference.  Times `updates` updates (after one
 * warm-up) of each candidate decomposition of this simulation's stage
 * kernels -- 2D: rows per wave x waves per workgroup; 3D first order: form,
 * cells per lane and planes per chunk; 3D Roe family: planes per chunk --
 * restores the state, clock and primitives after every trial, and keeps the
 * fastest.  Every decomposition gives the same bits.  The winner is kept per
 * grid shape for the process and, when HYDRO_TUNE_FILE names a file,
 * appended there; simulations created later with the same shape start from
 * it (unless a decomposition switch is set in the environment).  1D grids
 * have nothing to tune (returns the current choice).  params_out as
 * hydro_sim_decomposition (may be NULL). */
int hydro_sim_autotune(hydro_sim* sim, int updates, int* params_out, double* ms_per_update_out);

/* Device HIP stream of the sim (hipStream_t as void*), for callers that want
 * to order their own work against it. */
void* hydro_sim_stream(hydro_sim* sim);

/* Measurement hook (no counterpart in the reference, whose event profiling
 * is commented out: src/Solver/Solver.cpp:537-544).  While enabled, every
 * phase of an update (below) is bracketed by HIP events on the sim's stream,
 * and run() launches kernel by kernel instead of replaying graphs. */
int hydro_sim_set_profiling(hydro_sim* sim, int enable);
/* Average duration in ms of the fused stage launches timed since profiling
 * was enabled, and their count.  Synchronises the stream. */
int hydro_sim_stage_timing(hydro_sim* sim, double* avg_ms, long long* launches);

/* Phases of an update timed while profiling is enabled (the reference's
 * EventProfileEntry per kernel, include/HydroGPU/Solver/Solver.h:24-31, whose
 * readout src/Solver/Solver.cpp:537-544 keeps commented out):
 *   BOUNDARY  ghost fill (and the deferred ghost copy it needs)
 *   CFL       CFL kernel + dt tail of calc_timestep, or of an update whose
 *             boundary and CFL are separate launches
 *   HEAD      boundary and CFL march in one launch (an update's head)
 *   DT_TAIL   the dt tail after HEAD, or the clock of step(dt)
 *   STAGES    the fused RK stage launches
 *   CON2PRIM  SRHD constrain + primitive recovery
 *   PRESSURE  EulerBurgers pressure, momentum-stage and work launches
 *   GHOSTS    deferred ghost copies outside a boundary
 *   TOTALS    the conserved-totals reduction
 *   GRAVITY   the gravity source pass and the potential's relaxation sweeps
 * With profiling enabled each phase is also a roctx range
 * ("hydro:<phase>") when the ROCm tracer library can be loaded, so that
 * rocprofv3 --marker-trace groups kernels by phase. */
enum {
  HYDRO_PHASE_BOUNDARY = 0, HYDRO_PHASE_CFL, HYDRO_PHASE_HEAD, HYDRO_PHASE_DT_TAIL, HYDRO_PHASE_STAGES,
  HYDRO_PHASE_CON2PRIM, HYDRO_PHASE_PRESSURE, HYDRO_PHASE_GHOSTS, HYDRO_PHASE_TOTALS, HYDRO_PHASE_GRAVITY,
  HYDRO_NUM_PHASES
};
/* Total device time in ms (total_ms[HYDRO_NUM_PHASES]) and the number of
 * timed intervals (launches[HYDRO_NUM_PHASES], may be NULL) per phase since
 * profiling was enabled.  Synchronises the stream. */
int hydro_sim_phase_timing(hydro_sim* sim, double* total_ms, long long* launches);

/* ---- generic stage evaluation (host-side integrators) -------------------
 *
 * The reference's Integrator::integrate(dt, callback)
 * (include/HydroGPU/Integrator/Integrator.h:13-17) keeps the Runge-Kutta
 * combination on the host (RungeKutta.h:59-125, ForwardEuler.cpp:8-32) and
 * calls back into the solver for each derivative (Roe::step / calcDeriv,
 * src/Solver/Roe.cpp:94-112; HLL::step, src/Solver/HLL.cpp:234-241).  These
 * entry points give such an integrator the same pieces: a derivative of the
 * current state, and the buffer operations it combines stages with.  Buffers
 * are device arrays in the simulation's storage layout (N state planes of
 * `volume` elements of hydro_sim_element_size bytes, ghost cells included),
 * owned by the simulation.  Plain storage only (not HYDRO_FLAG_KAHAN).
 * hydro_sim_step / update / run remain the fused path for the same solvers.
 * Order: hydro_sim_boundary, hydro_sim_calc_timestep, hydro_sim_init_step,
 * then the integrator's stages (calc_deriv, copy, zero, mult_add on the
 * state buffer), then hydro_sim_finish_step. */

/* Bytes per element of the simulation's buffers: 8 (FP64) or 4 (FP32_MIXED). */
size_t hydro_sim_element_size(const hydro_sim* sim);

/* The coefficients alphas(i,k), betas(i,k) of an integrator
 * (include/HydroGPU/Integrator/RungeKutta.h:129-294; ForwardEuler: 1 stage,
 * alpha = beta = 1): *order stages, alpha[4*i+k], beta[4*i+k] (4x4 each). */
int hydro_integrator_tableau(int integrator, int* order, double* alpha, double* beta);

/* A zeroed device buffer of one state's size (Solver::CL::alloc); released
 * with hydro_sim_buffer_destroy or when the simulation is destroyed. */
int hydro_sim_buffer_create(hydro_sim* sim, void** buf);
int hydro_sim_buffer_destroy(hydro_sim* sim, void* buf);

/* The simulation's state buffer itself (Solver::stateBuffer).  Valid until
 * the next fused update (hydro_sim_step / update / run), which may move the
 * state to the other ping-pong buffer. */
int hydro_sim_state_buffer(hydro_sim* sim, void** buf);

/* Solver::initStep for a host-side step: freezes the start-of-step data
 * (Roe eigenbasis, HLL wave speeds, SRHD primitives) of the current state. */
int hydro_sim_init_step(hydro_sim* sim);

/* Roe::calcDeriv / HLL::calcDeriv (src/Solver/Roe.cpp:101-112,
 * res/CalcFluxDeriv.cl:7-44): deriv += -(flux differences)/dx over the
 * interior cells for the state buffer's current contents, with the frozen
 * data of hydro_sim_init_step and the full-step dt in the limiter
 * (res/Roe.cl:206, 292).  Accumulates, as the reference's kernel does: the
 * integrator zeroes `deriv` first.  `deriv` must come from
 * hydro_sim_buffer_create. */
int hydro_sim_calc_deriv(hydro_sim* sim, double dt, void* deriv);

/* enqueueCopyBuffer, Solver::CL::zero (src/Solver/Solver.cpp:22-32) and
 * multAdd a = b + c * k (res/Common.cl:201-210) on the simulation's buffers
 * (created ones or the state buffer), in stream order. */
int hydro_sim_buffer_copy(hydro_sim* sim, void* dst, const void* src);
int hydro_sim_buffer_zero(hydro_sim* sim, void* buf);
int hydro_sim_mult_add(hydro_sim* sim, void* a, const void* b, const void* c, double k);

/* Ends a host-side step: advances the simulated time by dt and the step
 * count, and runs the solver's tail (SRHD constrainState + updatePrimitives,
 * src/Solver/SRHDRoe.cpp:62-73). */
int hydro_sim_finish_step(hydro_sim* sim, double dt);

/* ---- many uneven simulations in one launch ------------------------------ */

/* A batch of independent Euler-Roe grids (1D or 2D, any mix of sizes) that
 * advance together: one persistent launch per RK stage covers every member,
 * dt is reduced per member.  All members share equation, solver, limiter,
 * integrator and precision; sizes, domains and boundaries may differ. */
int hydro_batch_create(hydro_ctx* ctx, const hydro_sim_desc* descs, int count, hydro_batch** out);
void hydro_batch_destroy(hydro_batch* b);
int hydro_batch_set_state(hydro_batch* b, int member, const double* aos, size_t n);
int hydro_batch_get_state(hydro_batch* b, int member, double* aos, size_t n);
/* nsteps updates of every member; use_graph != 0 captures one step in a
 * hipGraph and replays it. */
int hydro_batch_run(hydro_batch* b, int nsteps, int use_graph);
int hydro_batch_sync(hydro_batch* b);
int hydro_batch_member_dt(hydro_batch* b, int member, double* dt_out);
/* simulated time, completed steps and last dt of one member (synchronises) */
int hydro_batch_member_time(hydro_batch* b, int member, double* t, long long* steps, double* last_dt);
/* health flags of one member, as hydro_sim_health */
int hydro_batch_member_health(hydro_batch* b, int member, unsigned* flags_out);
/* total interior cells over all members */
long long hydro_batch_interior_cells(const hydro_batch* b);
/* Work decomposition: row-chunk height of the 2D members' tasks and strips
 * (60 columns each) per task of the 1D members.  No counterpart in the
 * reference (one program per grid there). */
int hydro_batch_set_decomposition(hydro_batch* b, int rows, int strips);
/* Autotuner: times candidate decompositions on the batch itself (the state,
 * clocks and buffer parity are restored afterwards), keeps the fastest and
 * reports it with its time per step. */
int hydro_batch_autotune(hydro_batch* b, int* rows_out, int* strips_out, double* ms_per_step_out);

#ifdef __cplusplus
}
#endif

#endif /* HYDRO_CAPI_H */
