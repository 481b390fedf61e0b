// lsm_internal.hpp - device-side mechanism tables and launch parameters shared by
// the host API (lsm_api.cpp) and the kernels (lsm_kernels.hip).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "lsm_bdf.hpp"

namespace lsm {

// Rows of the species coefficient table DevMech::sp ([NSP_ROWS][ns_pad], species
// along the fast axis so that lane k reads species k coalesced).  The blocks are
// the ones ThermodynamicsMap_CHEMKIN::SetCoefficients builds
// (maps/ThermodynamicsMap_CHEMKIN.hpp:263-343).
enum SpRow {
    SP_CP_HT = 0, SP_CP_LT = 5, SP_DH_HT = 10, SP_DH_LT = 16, SP_DS_HT = 22, SP_DS_LT = 28,
    SP_TMID = 34, SP_MW = 35, NSP_ROWS = 36
};

// Compact mechanism image that the team kernel copies into LDS once per
// workgroup (byte offsets from the image start; every block 16-byte aligned).
//   sp   double [LSP_ROWS][ns_pad]  rows 0-6 high-T a1..a7, 7-13 low-T a1..a7, 14 T_mid, 15 MW
//   P    double [nr][4]             lnA, beta, E/R, change of moles (device reaction order);
//                                   a PLOG reaction holds its PLOG slot in lnA, an
//                                   explicitly reversible one its rvp slot in the
//                                   change of moles
//   W    uint   [nr][4]             info, reactant bytes, product bytes, general list:
//        info: bit0 reversible, bit1 negative A, bit2 fixed-format record, bits 3/4
//              non-elementary forward/reverse side, bit5 explicit reverse Arrhenius,
//              bit6 chemically activated (CABR) fall-off entry, bit7 PLOG (beta field 0)
//              or Chebyshev (beta field 1) with its slot in the lnA field, bits 8-15
//              third-body slot+1, bits 16-23 fall-off slot+1, bits 24-31 number of
//              explicit forward orders (their entries follow the products);
//        fixed-format record: up to 3 reactant and 3 product species, integer orders as
//              repeated species, unused slots = LdsImage::dummy (c = 1, g = 0): 63
//              (LSM_DUMMY_SP) when ns <= 63, else 255;
//        general list: first entry (bits 0-15), reactants (16-23), products (24-31)
//   ent  uint [nent]                species (bits 0-11) | coefficient code << 12; code c < 16 means
//                                   nu = c/2, otherwise nu = nuv[c - 16]
//   nuv  double []                  coefficients that are not multiples of 1/2
//   tbp  uint [ntb+1], tbs uint [], tbe double []   third-body lists (eff - 1)
//   fop  double [nfo][8]            lnA_inf, beta_inf, E/R_inf, a, T***, T*, T**, type
//   fos  int [nfo] single collider or -1; fol uint [nfo+1]; fos2 uint []; foe double []
//   rvp  double [nrev][4]           explicit reverse lnA, beta, E/R, 0
//   plp  uint [npl+1], pll double [levels][4]   PLOG levels: ln p, lnA, beta, E/R
//   chp  uint [nch+1], chd double []            Chebyshev: N, M, Tmin, Tmax, log10 Pmin,
//                                               log10 Pmax, conv, a[N*M]
enum { LSM_DUMMY_SP = 63 };
enum { LSP_ROWS = 16 };
struct LdsImage {
    const void *src;
    int bytes;
    int o_sp, o_P, o_W, o_ent, o_nuv, o_tbp, o_tbs, o_tbe, o_fop, o_fos, o_fol, o_fos2, o_foe;
    int o_rvp;     // [nrev][4] explicit reverse lnA, beta, E/R, 0
    int o_plp;     // [npl+1] first level of each PLOG reaction
    int o_pll;     // [levels][4] ln p, lnA, beta, E/R
    int o_chp;     // [nch+1] first double of each Chebyshev reaction in chd
    int o_chd;     // per Chebyshev reaction: N, M, Tmin, Tmax, log10 Pmin, log10 Pmax, conv, a[N*M]
    int team_ok;   // the team kernel can use the image (ns <= 63)
    int wide_ok;   // the wide kernel can use the image (ns <= 254)
    int dummy;     // unused slot of fixed-format records: 63 if ns <= 63, else 255
    int simple0;   // first of the reactions, all at the end, that are fixed-format records with no
                   // collider, tabulated rate or explicit reverse (the team kernel's paired loop);
                   // nr when there are none or the layout is the wide kernel's
};

// Reactions are renumbered on upload: fall-off first, then third-body, then
// elementary, so that the lanes that take the extra fall-off and third-body work
// share as few passes of the reaction loop as possible (lsm_api.cpp).
struct DevMech {
    int ns, nr, ne, ntb, nfo, ns_pad;
    int xk;                                  // 1: explicitly reversible, chemically activated or PLOG
                                             // reactions present (kernels built with those forms)
    const double *sp;                        // [NSP_ROWS][ns_pad]
    const double *rA, *rB, *rE, *rdn;        // [nr] lnA, beta, E/R, change of moles
    const int *rinfo;                        // [nr] LSM_RXN_* bits
    const int *rtb, *rfo;                    // [nr] third-body / fall-off slot or -1
    const int *fw_ptr, *fw_sp;               // reactants sorted by species
    const double *fw_nu;
    const int *bw_ptr, *bw_sp;               // products sorted by species
    const double *bw_nu;
    const int *rx_ptr, *rx_sp;               // merged signed list (dH, dS)
    const double *rx_nu;
    const int *tb_ptr, *tb_sp;               // third-body efficiencies (eff - 1)
    const double *tb_eff;
    const int *fo_type, *fo_single;          // fall-off
    const double *fo_par;                    // [nfo][8] lnA_inf beta_inf EoR_inf a T3 T1 T2 0
    const int *fo_ptr, *fo_sp;
    const double *fo_eff;
    const int *f_ptr, *f_rx;                 // species-major formation list
    const double *f_nu;                      // signed coefficient
    const double *rv;                        // [nr][3] explicit reverse lnA, beta, E/R (zeros elsewhere)
    const int *rpl;                          // [nr] PLOG slot or -1
    const int *pl_ptr;                       // [npl+1] first level of each PLOG reaction
    const double *pl_lev;                    // [levels][4] ln p, lnA, beta, E/R
    const int *rch;                          // [nr] Chebyshev slot or -1
    const int *ch_ptr;                       // [nch+1] first double of each Chebyshev reaction
    const double *ch_dat;                    // N, M, Tmin, Tmax, log10 Pmin, log10 Pmax, conv, a[N*M]
    LdsImage img;
};

// Transport fits and mixture-rule constants (TransportPropertiesMap_CHEMKIN) on the device.
struct DevTransport {
    int ns, nteta, model;                    // model: LSM_VISCOSITY_*
    const double *mw;                        // [ns] kg/kmol
    const double *lam, *eta;                 // [ns][4] ln-T fits
    const double *gam;                       // [ns(ns-1)/2][4] ln-T fits of ln D_jk at 1 bar, pairs j < k
    const double *wsup, *winf, *wr14;        // [ns(ns-1)/2] Wilke pair constants (CompleteInitialization)
    const double *ww, *wr;                   // [ns][ns], partner-major: Wilke weight and sqrt-ratio factor of
                                             // species k with partner q at [q ns + k] (0 on the diagonal)
    const int *teta_slot;                    // [ns] row of teta, or -1
    const double *teta;                      // [nteta][ns][4] T fits of the thermal diffusion ratios
    // physical-model options (lsm_set_transport_options)
    int dmodel;                              // LSM_DIFFUSIVITY_*
    int simplified, soret;                   // 0/1
    int soot_ref = -1;                       // reference species of the soot corrections, or -1
    const double *lewis;                     // [ns] Lewis numbers (LSM_DIFFUSIVITY_LEWIS_NUMBERS)
    const int *soot_slot;                    // [ns] row of soot_corr, or -1 (null when soot_ref < 0)
    const double *soot_corr;                 // [n_soot]
};

struct KParams {
    BdfConst bk;
    unsigned long long *phase_cycles;   // [12] per-phase cycle totals (builds with LSM_PHASE_TIMERS only)
    double pressure;
    double min_temperature;
    int energy;
    int full_pivoting;
    int jac_mode;                       // 0: forward differences (KernelDense), 1: analytic
    int refine;                         // wide kernel: most refinement sweeps after an FP32 solve
    unsigned long long *fp64_fallbacks; // wide kernel: count of factorisations redone in FP64 (may be null)
    int const_volume;                   // constant-volume reactor (BatchReactorHomogeneousConstantVolume)
    int compact;                        // Strang "compact" output: y_out holds RR [kg/m3/s] and RT [W/m3]
    const double *rho;                  // constant volume: per-cell density [kg/m3] of this call (may be null:
                                        // ideal-gas density of the initial state at `pressure`)
};

// Host-side launchers (lsm_kernels.hip).
int team_max_neq();
// LDS layout of a team workgroup: the mechanism image, then per wave team_wave_bytes(ne).
// Per-wave LDS of the team kernel: species scratch (c and g pairs, R, dR/dT, h, collider
// terms: TEAM_WAVE_FIXED doubles), then the Jacobian, ne columns of 65 doubles (lane k's row at
// offset k of each column; the 65th slot staggers the banks), rounded up to 16 bytes so that
// every wave's (c, g) pairs stay 16-byte aligned.
enum { TEAM_WAVE_FIXED = 384 };
__host__ __device__ inline size_t team_wave_bytes(int ne) {
    return ((size_t)8 * (TEAM_WAVE_FIXED + 65 * (size_t)ne) + 15) / 16 * 16;
}
// Waves per SIMD the team kernel's registers are capped for.  Every register-array size needs
// more than 256 registers (its LU rows live in registers): a tighter cap only moves values to
// scratch memory (776 bytes per lane at three waves per SIMD for the 16-wide array).
#define TEAM_WAVES_PER_SIMD(NMAX) 1
int team_waves_per_cu(const DevMech &m);
// Whether blocks_per_cu workgroups of wpb waves (each with its own copy of the mechanism image)
// fit in lds_cu bytes of LDS.
bool team_shape_fits(const DevMech &m, int wpb, int blocks_per_cu, size_t lds_cu);
// order (may be null): the cells in the order the queue hands them out.
int launch_team(const DevMech &m, const KParams &k, int64_t n, const double *dt, const double *yin,
                double *yout, int *status, double *q, int *counters, unsigned *queue, const int *order,
                int grid, int waves_per_block, void *stream);
// Cells per wavefront of the team kernel: 1 (one cell per wave), 2 (ns + 1 <= 32) or 4
// (ns + 1 <= 16); the packed variants run 2 waves per SIMD.
int team_max_cells_per_wave(const DevMech &m);
int team_waves_per_cu_cpw(const DevMech &m, int cpw);
bool team_shape_fits_cpw(const DevMech &m, int cpw, int wpb, int blocks_per_cu, size_t lds_cu);
int launch_team_cpw(const DevMech &m, const KParams &k, int64_t n, const double *dt, const double *yin,
                    double *yout, int *status, double *q, int *counters, unsigned *queue, const int *order, int grid,
                    int waves_per_block, int cells_per_wave, void *stream);
int launch_plain(const DevMech &m, const KParams &k, int64_t n, const double *dt, const double *yin,
                 double *yout, int *status, double *q, int *counters, double *work, int64_t slots,
                 void *stream);
size_t plain_work_doubles(const DevMech &m, int64_t slots);
// Reactor right-hand side f (SoA [ne][n]) and analytic Jacobian J ([n][ne][ne]) at
// the states yin; either output may be null.
int launch_properties(const DevMech &m, double p_default, int64_t n, const double *y, const double *p, double *out,
                      double *x, double *cpk, int num_cu, void *stream);
// Transport part of the property refresh: mu, lambda [n], dmix, teta [ns][n] (the last
// two may be null).  One wave per cell, lane k on species k (k + 64 ...).
int launch_transport(const DevMech &m, const DevTransport &t, double p_default, int64_t n, const double *y,
                     const double *p, double *mu, double *lam, double *dmix, double *teta, int num_cu, void *stream);
int launch_equations(const DevMech &m, const KParams &k, int64_t n, const double *yin, double *f, double *J,
                     int grid, int waves_per_block, void *stream);
int launch_reset_queue(unsigned *queue, void *stream);
int launch_order_keys(const int *counters, int64_t n, unsigned *keys, int *vals, void *stream);
int launch_fill(double *p, int64_t n, double v, void *stream);
int launch_gather_cells(const double *y, const double *dt, int64_t n, int ne, int64_t m, int64_t stride, double *ys,
                        double *dts, void *stream);

// Wide kernel: one cell per 256-thread workgroup (ns + 1 <= WIDE_MAX_NEQ), thread e owns
// element e, Newton matrix factorised in FP32 in LDS, Jacobian in an HBM workspace
// ([ne][ne] doubles per workgroup), mechanism image read from global memory.
enum { WIDE_THREADS = 256, WIDE_MAX_NEQ = 176 };
size_t wide_lds_bytes(const DevMech &m);
size_t wide_jac_doubles(const DevMech &m);
int launch_wide(const DevMech &m, const KParams &k, int64_t n, const double *dt, const double *yin, double *yout,
                int *status, double *q, int *counters, unsigned *queue, const int *order, int grid, double *jws,
                void *stream);

}  // namespace lsm

struct lsm_handle;
// Device bytes a handle holds (tables, workspaces), for the serving budget.
size_t lsm_handle_device_bytes(const lsm_handle *h);
int lsm_handle_neq(const lsm_handle *h);
// Default reactor pressure of the handle (lsm_reactor_options::pressure).
double lsm_handle_pressure(const lsm_handle *h);
// Sets the thread's last-error text and returns code.
int lsm_fail(int code, const char *msg);
