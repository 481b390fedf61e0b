/*
 * lsm_oracle.h - CPU restatement of laminarSMOKE's operator-splitting reaction
 * step, used ONLY as a test oracle and as the CPU baseline of the benchmark.
 * Product code never links, loads or calls it.
 *
 * Reference files restated (paths relative to the laminarSMOKE repository;
 * OS/ = solvers/openSMOKEpp4laminarSMOKE/, LS/ = solvers/laminarSMOKE/):
 *   LS/unsteady/chemistry_DI.H:76-331                 per-cell driver
 *   LS/unsteady/BatchReactorHomogeneousConstantPressure.H:175-247   RHS
 *   OS/maps/ThermodynamicsMap_CHEMKIN.hpp:263-512     NASA-7
 *   OS/maps/KineticsMap_CHEMKIN.hpp:271-297,1181-1768,1956-1964   kinetics
 *   OS/maps/StoichiometricMap.hpp:512-730,916-1001    stoichiometry
 *   OS/math/native-ode-solvers/MultiValueSolver.hpp:36-1355,
 *     MethodGear.hpp:36-594, KernelDense.hpp:36-266   BDF (Gear) + FD Jacobian + LU
 *   OS/math/OpenSMOKEUtilities.hpp:339-345,689-701    UDot, ErrorControl
 *   LS/utilities.H:153-203                            normalizeMassFractions
 * Eigen 3.3.7 PartialPivLU / FullPivLU (not vendored) are restated as textbook
 * row-pivoted and rook-free complete-pivoting Doolittle LU.
 */
#ifndef LSM_ORACLE_H
#define LSM_ORACLE_H

#include <stdint.h>
#include "../include/lsm_mechanism.h"
#include "../include/lsm_transport.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LSM_ORACLE_NCOUNTERS 12

typedef struct lsm_oracle_options {
    double rel_tol;            /* OdeHomogeneous relTolerance                  */
    double abs_tol;            /* OdeHomogeneous absTolerance                  */
    int32_t full_pivoting;     /* OdeHomogeneous fullPivoting                  */
    int32_t max_steps;         /* <=0: 500000 (MultiValueSolver.hpp:51)        */
    double pressure;           /* thermodynamicPressure [Pa]                   */
    double min_temperature;    /* minTemperature4Chemistry [K]                 */
    int32_t energy_equation;   /* energyEquation                               */
    double mass_fractions_tol; /* massFractionsTol (warning threshold)         */
    double t0;                 /* start time of the chemistry step             */
    int32_t constant_volume;   /* constPressureBatchReactor false: constant-volume
                                  reactor (BatchReactorHomogeneousConstantVolume) */
    int32_t strang_compact;    /* strangAlgorithm STRANG_COMPACT: y_out holds the
                                  sources RR [kg/m3/s] and RT [W/m3]
                                  (chemistry_DI.H:236-282)                     */
    const double *rho;         /* constant volume: per-cell density rho0 [kg/m3]
                                  (rhoCells); NULL: ideal-gas density of the
                                  initial state at `pressure`                  */
    double jac_increment_scale; /* test option: factor on the forward-difference
                                  increments of the numerical Jacobian (<= 0 or 1:
                                  the reference's increments).  Other values give
                                  an equally valid Jacobian approximation, which
                                  measures how the reference's result depends on
                                  its Newton matrix.                          */
    int32_t reverse_sums;      /* test option: 1 sums each species' formation
                                  rate over its reactions in reverse order (0:
                                  the reference's order).  An equally valid
                                  rounding of the same sums, which measures how
                                  the reference's result depends on the order of
                                  its floating-point additions.               */
} lsm_oracle_options;

typedef void (*lsm_oracle_rhs_fn)(void *ctx, const double *y, double *f);

typedef struct lsm_oracle_bdf_options {
    double rel_tol, abs_tol;
    int32_t full_pivoting;
    int32_t max_steps;
    int32_t use_min, use_max;  /* SetMinimumValues / SetMaximumValues given    */
    const double *ymin, *ymax;
} lsm_oracle_bdf_options;

/* NASA-7 species properties at T (ThermodynamicsMap_CHEMKIN.hpp:436-496). */
void lsm_oracle_species_thermo(const lsm_mechanism *m, double T,
                               double *cp_R, double *h_RT, double *s_R);

/* Kinetics at (T, c): forward constants kf (after the sign flip), 1/Kc (0 for
 * irreversible reactions), effective constants kmod (third body, fall-off),
 * net rates r [kmol/m3/s] and formation rates R [kmol/m3/s]. Any output may be NULL. */
void lsm_oracle_kinetics(const lsm_mechanism *m, double T, const double *c,
                         double *kf, double *uKeq, double *kmod, double *r, double *R);
/* The same with the map set to pressure P [Pa] (it only enters PLOG constants;
 * lsm_oracle_kinetics uses 1 atm).  For explicitly reversible reactions uKeq
 * holds k_rev/k. */
void lsm_oracle_kinetics_p(const lsm_mechanism *m, double T, double P, const double *c,
                           double *kf, double *uKeq, double *kmod, double *r, double *R);

/* Transport properties of one cell at (T, P) and mole fractions x[ns]: the map's
 * species fits and mixture rules as properties.H:101-147 calls them
 * (TransportPropertiesMap_CHEMKIN.hpp:690-1117).  mu [kg/m/s], lambda [W/m/K],
 * dmix[ns] mixture-averaged diffusion coefficients [m2/s], teta[ns] thermal
 * diffusion ratios (0 for species without them); any output may be NULL. */
void lsm_oracle_transport(const lsm_transport *tr, double T, double P, const double *x, double *mu,
                          double *lambda, double *dmix, double *teta);

/* One cell of laminarSMOKE's property refresh with its physical-model options
 * (solvers/laminarSMOKE/properties.H:79-192): mixture mu and lambda (fitted, or the
 * simplified power law), diffusion coefficients (mixture-averaged or from Lewis numbers,
 * then the soot corrections), and theta or, with o->soret, Dsoret.  y: [ns+1] (Y, T).
 * o may be NULL (the defaults).  dmix / teta may be NULL. */
void lsm_oracle_refresh_cell(const lsm_mechanism *m, const lsm_transport *tr, const lsm_transport_options *o,
                             double P, const double *y, double *mu, double *lambda, double *dmix, double *teta);

/* Constant-pressure RHS (BatchReactorHomogeneousConstantPressure::Equations);
 * y = [Y_0..Y_{ns-1}, T], returns QR [J/m3/s]. */
double lsm_oracle_rhs(const lsm_mechanism *m, double P, int energy, const double *y, double *f);

/* Constant-volume RHS (BatchReactorHomogeneousConstantVolume::Equations,
 * BatchReactorHomogeneousConstantVolume.H:150-199) at density rho0 [kg/m3];
 * returns QR [J/m3/s]. */
double lsm_oracle_rhs_cv(const lsm_mechanism *m, double rho0, int energy, const double *y, double *f);

/* Dense LU, row-major n x n in place; rp/cp row/column permutations (cp may be
 * NULL for partial pivoting).  Returns 0, or k+1 if pivot k was exactly zero. */
int lsm_oracle_lu_factor(int n, double *A, int32_t *rp, int32_t *cp, int full);
void lsm_oracle_lu_solve(int n, const double *LU, const int32_t *rp, const int32_t *cp,
                         double *b, int full);

/* Generic Gear/BDF solve of dy/dt = f(y) from t0 to tOut (fresh solver).
 * Returns the OdeStatus; yout gets the solution (y0 on failure). */
int lsm_oracle_bdf(int n, lsm_oracle_rhs_fn f, void *ctx, const lsm_oracle_bdf_options *o,
                   double t0, const double *y0, double tOut, double *yout, int32_t *counters);

/* One cell of the chemistry step, AoS y = [Y, T]; returns the OdeStatus
 * (0 when the cell is below min_temperature and not integrated). */
int lsm_oracle_integrate_cell(const lsm_mechanism *m, const lsm_oracle_options *o, double dt,
                              const double *y_in, double *y_out, double *q, int32_t *counters);

/* The chemistry step over n cells, SoA y[i*n + cell] (i = 0..ns-1 species,
 * i = ns temperature), n_threads >= 1 worker threads with 64-cell chunks. */
void lsm_oracle_chemistry_step(const lsm_mechanism *m, const lsm_oracle_options *o,
                               int64_t n, const double *dt, const double *y_in, double *y_out,
                               int32_t *status, double *q, int32_t *counters, int n_threads);

#ifdef __cplusplus
}
#endif
#endif
