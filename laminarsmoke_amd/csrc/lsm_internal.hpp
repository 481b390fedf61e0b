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

// Reactions are renumbered on upload: elementary first, then third-body, then
// fall-off, so that the wavefront lanes that take the extra third-body and
// fall-off work are contiguous.  rperm maps device order -> mechanism order.
struct DevMech {
    int ns, nr, ne, ntb, nfo, ns_pad;
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
};

struct KParams {
    BdfConst bk;
    double pressure;
    double min_temperature;
    int energy;
    int full_pivoting;
};

// Host-side launchers (lsm_kernels.hip).
int team_max_neq();
size_t team_lds_bytes(const DevMech &m);
int launch_team(const DevMech &m, const KParams &k, int64_t n, const double *dt, const double *yin,
                double *yout, int *status, double *q, int *counters, unsigned *queue, int grid,
                void *stream);
int launch_plain(const DevMech &m, const KParams &k, int64_t n, const double *dt, const double *yin,
                 double *yout, int *status, double *q, int *counters, double *work, int64_t slots,
                 void *stream);
size_t plain_work_doubles(const DevMech &m, int64_t slots);
int launch_reset_queue(unsigned *queue, void *stream);

}  // namespace lsm
