/*
 * laminarsmoke_amd.h - C-ABI of the MI355X chemistry-step integrator.
 *
 * Drop-in boundary for laminarSMOKE's operator-splitting reaction step: the
 * per-cell constant-pressure batch-reactor integration that
 * solvers/laminarSMOKE/unsteady/chemistry_DI.H:76-331 performs serially with
 * OdeSMOKE::MultiValueSolver<MethodGear<KernelDense<...>>> becomes one batched
 * call.  All functions return 0 (LSM_OK) or a negative LSM_E_* code; the text of
 * the last error is available from lsm_last_error().  No function aborts or
 * exits the process (the reference's FatalErrorMessage does, MultiValueSolver.hpp:352-359).
 *
 * Threading: a handle may be used from one host thread at a time; different
 * handles are independent.  Device entry points enqueue work on the caller's
 * HIP stream and return without synchronising.
 */
#ifndef LAMINARSMOKE_AMD_H
#define LAMINARSMOKE_AMD_H

#include <stddef.h>
#include <stdint.h>

#include "lsm_mechanism.h"
#include "lsm_transport.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * ODE solver selection and parameters.
 * Mirrors OpenSMOKE::ODE_Parameters (math/external-ode-solvers/ODE_Parameters.h:52-160)
 * with one new integrator value, LSM_ODE_OPENSMOKE_GPU.
 * ------------------------------------------------------------------------- */
enum lsm_ode_integrator {              /* ODE_Parameters::ODE_INTEGRATOR (ODE_Parameters.h:52-55) */
    LSM_ODE_OPENSMOKE = 0,
    LSM_ODE_BZZODE = 1,
    LSM_ODE_CVODE = 2,
    LSM_ODE_DVODE = 3,
    LSM_ODE_DASPK = 4,
    LSM_ODE_DLSODE = 5,
    LSM_ODE_DLSODA = 6,
    LSM_ODE_RADAU5 = 7,
    LSM_ODE_MEBDF = 8,
    LSM_ODE_CHEMEQ2 = 9,
    LSM_ODE_OPENSMOKE_GPU = 100        /* this library: OpenSMOKE Gear/BDF on the GPU */
};

typedef struct lsm_ode_parameters {
    int32_t type;                      /* lsm_ode_integrator                                   */
    double relative_tolerance;         /* default 100*FLT_EPSILON (ODE_Parameters.hpp:248)     */
    double absolute_tolerance;         /* default 1e-10                                        */
    double minimum_step;               /* -1: solver default                                   */
    double maximum_step;               /* -1: solver default                                   */
    double initial_step;               /* -1: solver default                                   */
    int32_t maximum_order;             /* -1: solver default (5)                               */
    int32_t maximum_number_of_steps;   /* -1: solver default (500000)                          */
    int32_t full_pivoting;             /* 0/1: FullPivLU (KernelDense.hpp:209-237).  The plain
                                          kernel pivots fully; the team and wide kernels
                                          always pivot by rows (partial pivoting), which changes
                                          the Newton solves only at rounding level           */
    char linear_algebra[32];           /* "Eigen" (the only value KernelDense accepts)         */
} lsm_ode_parameters;

/* Defaults of ODE_Parameters::ODE_Parameters() (ODE_Parameters.hpp:239-256), with
 * type = LSM_ODE_OPENSMOKE_GPU. */
int lsm_ode_parameters_default(lsm_ode_parameters *p);

/* Parse a parameter dictionary into *p (starting from the values already in *p).
 * Accepts both the laminarSMOKE OpenFOAM sub-dictionary form read at
 * solvers/laminarSMOKE/readOptions.H:474-530
 *     OdeHomogeneous { odeSolver "OpenSMOKE_GPU"; relTolerance 1e-7;
 *                      absTolerance 1e-12; maximumOrder 5; fullPivoting false; }
 * and the OpenSMOKE @-keyword form of ODE_Parameters::SetupFromDictionary
 * (ODE_Parameters.hpp:39-236, grammar ODE_Parameters.h:163-239)
 *     @OdeSolver OpenSMOKE_GPU; @RelativeTolerance 1e-7; @FullPivoting true; ...
 * Unknown solver names, unknown keys and malformed values are rejected with
 * LSM_E_PARSE (readOptions.H:495-504 aborts instead). */
int lsm_ode_parameters_parse(const char *text, lsm_ode_parameters *p);

/* ---------------------------------------------------------------------------
 * Reactor / cell-loop options (the variables chemistry_DI.H reads).
 * ------------------------------------------------------------------------- */
typedef struct lsm_reactor_options {
    double pressure;            /* default thermodynamicPressure [Pa] of the property refresh
                                   (p NULL) and the serving layer; the chemistry step takes its
                                   pressure per call (chemistry_DI.H:134)                      */
    double min_temperature;     /* minTemperature4Chemistry: cells with T <= it are not
                                   integrated (chemistry_DI.H:97)                              */
    int32_t energy_equation;    /* energyEquation (chemistry_DI.H:135, 231-233)               */
    double mass_fractions_tol;  /* massFractionsTol: |sum Y - 1| warning threshold
                                   (utilities.H:159)                                          */
    int32_t constant_volume;    /* !constPressureBatchReactor: constant-volume reactor
                                   (BatchReactorHomogeneousConstantVolume.H:150-199, selected
                                   at chemistry_DI.H:171-208); numerical Jacobian only        */
    int32_t strang_compact;     /* strangAlgorithm == STRANG_COMPACT: y_out receives the
                                   sources RR_k = rho (Y_k,out - Y_k,in)/dt [kg/m3/s] (rows
                                   0..ns-1) and RT = rho cp (T_out - T_in)/dt [W/m3] (row ns),
                                   rho and cp of the final state at `pressure`; zero when
                                   dt <= 1e-14 (chemistry_DI.H:236-282)                       */
} lsm_reactor_options;

int lsm_reactor_options_default(lsm_reactor_options *r);

/* ---------------------------------------------------------------------------
 * Integrator selection inside the GPU library.
 * ------------------------------------------------------------------------- */
enum lsm_method {
    LSM_METHOD_AUTO = 0,        /* team when ns+1 <= 64, else wide when the LU fits LDS, else plain */
    LSM_METHOD_PLAIN = 1,       /* one cell per thread, FP64, state in HBM                    */
    LSM_METHOD_TEAM = 2,        /* one cell per wavefront, persistent work queue, FP64        */
    LSM_METHOD_WIDE = 3         /* one cell per 256-thread workgroup, persistent work queue;
                                   FP64 rates, residuals and error control; the Newton matrix
                                   inverted in FP32 in LDS (blocked Gauss-Jordan with partial
                                   pivoting, MFMA rank-16 updates), each solve an FP32
                                   mat-vec with FP64 iterative refinement (ns+1 <= 176)        */
};

/* OdeStatus values (math/native-ode-solvers/OdeSolverUtilities.h:83-99); a cell
 * below min_temperature reports LSM_STATUS_SKIPPED. */
enum {
    LSM_STATUS_SKIPPED = 0,
    LSM_STATUS_CONTINUATION = 2,
    LSM_STATUS_MAX_NUMBER_OF_STEPS_REACHED = -1,
    LSM_STATUS_TOO_STRICT_TOLERANCES = -2,
    LSM_STATUS_MAX_NUMBER_ERRORTEST_FAILURES = -4,
    LSM_STATUS_MAX_NUMBER_CONVERGENCETEST_FAILURES = -5,
    LSM_STATUS_TOO_SMALL_STEP_SIZE = -6,
    LSM_STATUS_ILLEGAL_CONSTRAINTS = -9
};

/* Per-cell counters (int32, LSM_NCOUNTERS per cell, cell-major), mirroring
 * MethodGear.h:85-120 / MultiValueSolver.h:199-234 / KernelDense.h. */
enum {
    LSM_CNT_STEPS = 0,                  /* numberOfSteps()                      */
    LSM_CNT_FUNCTION_CALLS = 1,         /* numberOfFunctionCalls()              */
    LSM_CNT_FUNCTION_CALLS_JACOBIAN = 2,/* numberOfFunctionCallsForJacobian     */
    LSM_CNT_JACOBIANS = 3,              /* numberOfJacobianEvaluations()        */
    LSM_CNT_FACTORIZATIONS = 4,         /* numberOfMatrixFactorizations()       */
    LSM_CNT_LINEAR_SOLUTIONS = 5,       /* numberOfLinearSystemSolutions()      */
    LSM_CNT_CONVERGENCE_FAILURES = 6,   /* numberOfConvergenceFailure()         */
    LSM_CNT_ERROR_TEST_FAILURES = 7,    /* numberOfErrorCheckFailure()          */
    LSM_CNT_MAX_ORDER_USED = 8,         /* maxOrderUsed()                       */
    LSM_CNT_RHS_TOTAL = 9,              /* every RHS evaluation (incl. Jacobian) */
    LSM_CNT_CLAMPS = 10,                /* yMin/yMax constraint activations     */
    LSM_CNT_NEWTON_ITERATIONS = 11,     /* sum of Newton iterations             */
    LSM_NCOUNTERS = 12
};

/* Error codes. */
enum {
    LSM_OK = 0,
    LSM_E_INVALID = -1,         /* bad argument                                 */
    LSM_E_PARSE = -2,           /* dictionary parse error                       */
    LSM_E_UNSUPPORTED = -3,     /* option / mechanism not supported             */
    LSM_E_HIP = -4,             /* HIP runtime error                            */
    LSM_E_NOMEM = -5            /* allocation failed                            */
};

typedef struct lsm_handle lsm_handle;

/* Upload the mechanism tables to the current HIP device and bind the ODE
 * parameters (replaces the construction of the maps at readOptions.H:76-78, the
 * reactor at memoryAllocation.H:51 and the solver at memoryAllocation.H:58-62). */
int lsm_create(const lsm_mechanism *mech, const lsm_ode_parameters *ode,
               const lsm_reactor_options *reactor, lsm_handle **out);

int lsm_set_method(lsm_handle *h, int32_t method);
int lsm_destroy(lsm_handle *h);

/* Cells per wavefront of the team kernel (no counterpart in the reference, whose cell loop
 * integrates one cell at a time, chemistry_DI.H:92-328): 1 (default) gives each cell a
 * whole 64-lane wavefront; 2 (ns+1 <= 32) and 4 (ns+1 <= 16) split the wavefront into
 * segments of 32 or 16 lanes, one cell each, so that small mechanisms leave fewer lanes
 * idle.  The arithmetic per cell is the same in every packing; a wavefront's segments
 * branch apart where their cells' step histories differ.  lsm_autotune searches it.
 * LSM_E_UNSUPPORTED when the mechanism is too large for the packing. */
int lsm_set_cells_per_wave(lsm_handle *h, int32_t cells);
int lsm_get_cells_per_wave(lsm_handle *h, int32_t *cells);

/* Jacobian used by the Newton iteration of the team kernel.  NUMERICAL (default)
 * is the forward-difference Jacobian of KernelDense::NumericalJacobian
 * (math/multivalue-ode-solvers/KernelDense.hpp:160-183), ne extra RHS evaluations
 * per Jacobian, as the reference does for every OpenSMOKE integration.  ANALYTIC
 * builds dR/dc and dR/dT of the CHEMKIN kinetics in closed form (no extra RHS
 * evaluations); it changes the Newton matrix only, not the error control, so the
 * result stays within the integration tolerance of the reference's.  The plain
 * kernel always uses NUMERICAL. */
typedef enum lsm_jacobian_type {
    LSM_JACOBIAN_NUMERICAL = 0,
    LSM_JACOBIAN_ANALYTIC = 1
} lsm_jacobian_type;
int lsm_set_jacobian(lsm_handle *h, int32_t type);

/* Most iterative-refinement sweeps after each FP32 Newton solve of the wide kernel
 * (0-8, default 4): x += G32^-1 (b - G x), G32^-1 the FP32 inverse of the Newton
 * matrix, with the residual in FP64 from the FP64
 * Jacobian, until the correction's WRMS norm in the integration weights is below 1e-5
 * (of the tolerance).  A factorisation whose solves do not get there (a sweep that
 * does not halve the correction, or no sweeps left) is redone in FP64 for this and the
 * following solves.  0: the FP32 solve alone, unchecked.  The team and plain kernels
 * factorise in FP64 and ignore it. */
int lsm_set_refinement(lsm_handle *h, int32_t sweeps);

/* Number of wide-kernel factorisations redone in FP64 since lsm_create (synchronises the
 * device).  No counterpart in the reference, which factorises in FP64 only. */
int lsm_fp64_fallbacks(lsm_handle *h, uint64_t *count);

/* Chemistry step over n_cells cells, host pointers (the whole forAll loop of
 * chemistry_DI.H:92-328).
 *   t0: start time of the step, the t0 = t - dt of the CFD time step
 *       (Policy_MomentumTransportReaction.H:62) that the loop passes to
 *       SetInitialConditions(t0, y0) and Solve(t0 + DeltaTCells[celli])
 *       (chemistry_DI.H:138,157).  It is not a mere offset: it sets the bounds of the
 *       first step size (MultiValueSolver.hpp:649-685) and the step-too-small test
 *       t + h == t (:812-820), so stiff cells late in a run can end with status -6.
 *   pressure: thermodynamicPressure [Pa] of the reactors (SetReactor(P),
 *       chemistry_DI.H:134), > 0.
 *   y_in/y_out: SoA [ns+1][n_cells] (species rows, then T).  dt: per-cell integration
 *   time (DeltaTCells).  status (may be NULL): OdeStatus per cell.  q (may be NULL):
 *   heat release QR [W/m3] of integrated cells; entries of skipped cells are left
 *   untouched (QCells keeps its value).  counters (may be NULL):
 *   [n_cells][LSM_NCOUNTERS].  Synchronous. */
int lsm_integrate_batch(lsm_handle *h, int64_t n_cells, double t0, double pressure, const double *dt,
                        const double *y_in, double *y_out, int32_t *status, double *q, int32_t *counters);

/* Same contract on device pointers, enqueued on `stream` (a hipStream_t; NULL =
 * default stream).  y_in and y_out may alias.  Not synchronising. */
int lsm_integrate_batch_device(lsm_handle *h, int64_t n_cells, double t0, double pressure, const double *d_dt,
                               const double *d_y_in, double *d_y_out, int32_t *d_status,
                               double *d_q, int32_t *d_counters, void *stream);

/* lsm_integrate_batch_device with a visiting order: the persistent kernels (team,
 * wide) hand out cells in the order d_order[0..n_cells) (a permutation of the cell
 * indices; entries outside [0, n_cells) are skipped).  Results do not depend on the
 * order; the makespan does: expensive cells first shortens the tail.  The plain
 * kernel ignores the order.  No counterpart in the reference, whose cell loop is
 * serial (chemistry_DI.H:92). */
int lsm_integrate_batch_device_ordered(lsm_handle *h, int64_t n_cells, double t0, double pressure,
                                       const double *d_dt, const double *d_y_in, double *d_y_out,
                                       int32_t *d_status, double *d_q, int32_t *d_counters,
                                       const int32_t *d_order, void *stream);

/* Constant-volume variant: as lsm_integrate_batch / _device_ordered (d_order may be
 * NULL), with the per-cell density rho0 [kg/m3] of the constant-volume reactor,
 * rhoCells in BatchReactorHomogeneousConstantVolume::SetReactor(V, P, rho)
 * (chemistry_DI.H:173).  rho NULL: the ideal-gas density of each cell's initial state
 * at `pressure`.  Constant-pressure handles ignore rho. */
int lsm_integrate_batch_rho(lsm_handle *h, int64_t n_cells, double t0, double pressure, const double *dt,
                            const double *rho, const double *y_in, double *y_out, int32_t *status, double *q,
                            int32_t *counters);
int lsm_integrate_batch_device_rho(lsm_handle *h, int64_t n_cells, double t0, double pressure,
                                   const double *d_dt, const double *d_rho, const double *d_y_in,
                                   double *d_y_out, int32_t *d_status, double *d_q, int32_t *d_counters,
                                   const int32_t *d_order, void *stream);

/* lsm_integrate_batch_device_ordered (d_order may be NULL) that allocates no memory:
 * no counterpart in the reference; for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The scratch
 * its kernels need besides inputs and outputs (the work-queue counter; the wide
 * kernel's per-workgroup Jacobian and FP64 LU; the plain kernel's per-cell
 * workspaces) is a static device array of LSM_NOALLOC_SCRATCH_BYTES per GPU, held
 * until the process ends; the wide and plain kernels run on as many workgroups / cell
 * slots as it holds (same results, possibly fewer in flight), and a mechanism whose
 * single workspace exceeds it is rejected (LSM_E_INVALID).  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams, other
 * threads, other handles) race and must be serialised by the caller.  The queue
 * counter is reset by a kernel on `stream` at every call.  FP64 refactorisations of
 * the wide kernel are not added to lsm_fp64_fallbacks.  The current device must be
 * the handle's; constant-volume handles are rejected (they need per-cell densities). */
enum { LSM_NOALLOC_SCRATCH_BYTES = 134217728 };   /* 128 MiB */
int lsm_integrate_batch_device_noalloc(lsm_handle *h, int64_t n_cells, double t0, double pressure,
                                       const double *d_dt, const double *d_y_in, double *d_y_out,
                                       int32_t *d_status, double *d_q, int32_t *d_counters,
                                       const int32_t *d_order, void *stream);

/* Visiting order from the per-cell counters of a previous call on the same cells
 * (the cpuChemistry field of chemistry_DI.H:170 plays this role in the reference):
 * cells sorted by BDF steps, most first (device radix sort).  Asynchronous on
 * `stream`; buffers are kept in the handle. */
int lsm_cost_order(lsm_handle *h, int64_t n_cells, const int32_t *d_counters, int32_t *d_order,
                   void *stream);

/* Launch-shape autotuner for the team kernel: times cells per wavefront (1, 2, 4 as the
 * mechanism allows, lsm_set_cells_per_wave) x waves per workgroup (the workgroup shares
 * one LDS copy of the mechanism) x workgroups per CU on a strided sample of `sample`
 * cells (0: 65536) of the given device batch, keeps the fastest in the handle.  t0 and pressure: those of the steps to come (lsm_integrate_batch).
 * result (may be NULL): {waves per workgroup, grid, best time in microseconds}.
 * Synchronous; allocates temporaries.  Other methods: no-op. */
int lsm_autotune(lsm_handle *h, int64_t n_cells, double t0, double pressure, const double *d_dt,
                 const double *d_y, int64_t sample, void *stream, int32_t *result);

/* Launch configuration in use: {method, waves per workgroup, grid, Jacobian type,
 * refinement sweeps}. */
int lsm_get_launch(lsm_handle *h, int32_t *out5);

/* Device bytes the handle holds now: mechanism tables, kernel workspaces, and the
 * buffers it grows on demand (host-call staging, plain-kernel workspace, cost-order
 * sort, transport tables).  No counterpart in the reference; for callers that
 * budget HBM (include/lsm_serve.h counts it). */
int lsm_device_bytes(lsm_handle *h, uint64_t *bytes);

/* Right-hand side and analytic Jacobian of the constant-pressure batch reactor at
 * n_cells states: BatchReactorHomogeneousConstantPressure_ODE_OpenSMOKE::GetEquations
 * (solvers/laminarSMOKE/unsteady/BatchReactorHomogeneousConstantPressure_ODE_Interface.H:96)
 * on a batch, plus the Jacobian that GetAnalyticalJacobian (same file, :137) leaves
 * empty.  y, f: SoA [ns+1][n_cells]; jac: [n_cells][ns+1][ns+1] row-major,
 * jac[c][i][j] = d f_i / d y_j.  f or jac may be NULL.  Host pointers, synchronous;
 * the _device variant takes device pointers and a hipStream_t and does not
 * synchronise.  Needs ns+1 <= 64. */
int lsm_equations(lsm_handle *h, int64_t n_cells, const double *y, double *f, double *jac);
int lsm_equations_device(lsm_handle *h, int64_t n_cells, const double *d_y, double *d_f, double *d_jac,
                         void *stream);

/* Thermodynamic part of the per-cell property refresh that follows the chemistry
 * step in every PIMPLE iteration (solvers/laminarSMOKE/properties.H:41-131): mole
 * fractions X, MWmix, cTot = p/(RT), psi = cTot MWmix/p, h, cp, cv and the species
 * cp (CpSpecies, with massDiffusionInEnergyEquation), per cell.  The transport part of
 * that loop is lsm_transport_properties below.
 *   y: SoA [ns+1][n_cells] (Y, then T);  p: [n_cells] pressure, or NULL for the
 *   handle's reactor pressure;  props: [LSM_NPROPS][n_cells];  x, cp_species:
 *   [ns][n_cells] or NULL.  Host variant synchronous; _device takes device pointers
 *   and a hipStream_t and does not synchronise. */
enum {
    LSM_PROP_MW = 0,      /* kg/kmol                          */
    LSM_PROP_CTOT = 1,    /* kmol/m3                          */
    LSM_PROP_PSI = 2,     /* s2/m2 (rho = psi p)              */
    LSM_PROP_H = 3,       /* J/kg                             */
    LSM_PROP_CP = 4,      /* J/kg/K                           */
    LSM_PROP_CV = 5,      /* J/kg/K                           */
    LSM_NPROPS = 6
};
int lsm_properties(lsm_handle *h, int64_t n_cells, const double *y, const double *p, double *props, double *x,
                   double *cp_species);
int lsm_properties_device(lsm_handle *h, int64_t n_cells, const double *d_y, const double *d_p, double *d_props,
                          double *d_x, double *d_cp_species, void *stream);

/* Transport part of the same loop (properties.H:101-147): the mixture viscosity
 * (TransportPropertiesMap::DynamicViscosity), thermal conductivity
 * (ThermalConductivity), mixture-averaged diffusion coefficients
 * (MassDiffusionCoefficients) and thermal diffusion ratios (ThermalDiffusionRatios),
 * from the fits of include/lsm_transport.h
 * (maps/TransportPropertiesMap.hpp:41-71, maps/TransportPropertiesMap_CHEMKIN.hpp:690-1117).
 * lsm_set_transport copies the fits to the device (tr->ns must equal the mechanism's);
 * it may be called again to replace them.
 *   y: SoA [ns+1][n_cells] (Y, then T);  p: [n_cells] or NULL (the handle's pressure);
 *   mu [kg/m/s], lambda [W/m/K]: [n_cells];  dmix [m2/s], teta [-]: [ns][n_cells] or
 *   NULL (teta is 0 for species without thermal diffusion ratios).  Host variant
 *   synchronous; _device takes device pointers and a hipStream_t, no synchronisation. */
int lsm_set_transport(lsm_handle *h, const lsm_transport *tr);

/* The physical-model options of the same refresh (readOptions.H:94-181, 283-360;
 * properties.H:109-192): Lewis-number diffusivities, simplified transport properties,
 * Soret diffusion coefficients and soot-diffusivity corrections; include/lsm_transport.h
 * gives their meaning.  NULL restores the defaults (mixture-averaged diffusion, the
 * fitted mu and lambda, theta output, no corrections).  Takes effect for the following
 * lsm_transport_properties calls; may be called before or after lsm_set_transport. */
int lsm_set_transport_options(lsm_handle *h, const lsm_transport_options *o);
int lsm_transport_properties(lsm_handle *h, int64_t n_cells, const double *y, const double *p, double *mu,
                             double *lambda, double *dmix, double *teta);
int lsm_transport_properties_device(lsm_handle *h, int64_t n_cells, const double *d_y, const double *d_p,
                                    double *d_mu, double *d_lambda, double *d_dmix, double *d_teta, void *stream);

/* Text of the last error of this thread ("" if none). */
const char *lsm_last_error(void);

/* Library version string. */
const char *lsm_version(void);

#ifdef __cplusplus
}
#endif

#endif /* LAMINARSMOKE_AMD_H */
