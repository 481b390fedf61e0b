/*
 * lsm_transport.h - flat, read-only transport fits of a gas mixture, shared by the
 * HIP library (laminarsmoke_amd) and the CPU oracle.
 *
 * The tables are what OpenSMOKE++'s TransportPropertiesMap_CHEMKIN holds after
 * reading the <Transport type="CHEMKIN"> block of kinetics.xml
 * (maps/TransportPropertiesMap_CHEMKIN.hpp:288-338, 340-403), in SI units:
 *
 *   mw[k]                 molecular weight used by the mixture rules [kg/kmol]
 *   lambda_fit[4k+i]      ln lambda_k = sum_i a_i (ln T)^i, lambda in W/m/K    (:690-727)
 *   eta_fit[4k+i]         ln eta_k    = sum_i a_i (ln T)^i, eta in kg/m/s      (:729-766)
 *   gamma_fit[4q+i]       pair q = (j, k), j < k, in the order j = 0..ns-1,
 *                         k = j+1..ns-1: ln(D_jk at 1 bar) = sum_i a_i (ln T)^i, D in
 *                         m2/s; the map keeps 1/D_jk = (P / 1e5 Pa) / exp(...)  (:798-859)
 *   teta_sp[m]            the species with thermal diffusion ratios (MW <= 20,
 *                         PhysicalConstants.h:44), teta_fit[4(m ns + j) + i]:
 *                         theta_mj = sum_i a_i T^i                              (:768-796)
 *   viscosity_model       LSM_VISCOSITY_* (ImportViscosityModelFromXMLFile, :634-651)
 *
 * All indices are 0-based.  The arrays are owned by the caller and are only read
 * during lsm_set_transport()/oracle calls.
 */
#ifndef LSM_TRANSPORT_H
#define LSM_TRANSPORT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    LSM_VISCOSITY_WILKE = 0,         /* Wilke as modified by Bird et al. (:959-1004, default) */
    LSM_VISCOSITY_HERNING = 1,       /* Herning-Zipperer (:1006-1018)                        */
    LSM_VISCOSITY_MATHUR_SAXENA = 2  /* 0.5 (sum x eta + 1/sum x/eta) (:1021-1027)           */
};

typedef struct lsm_transport {
    int32_t ns;                  /* number of species (the mechanism's)              */
    const double *mw;            /* [ns]                                             */
    const double *lambda_fit;    /* [ns*4]                                           */
    const double *eta_fit;       /* [ns*4]                                           */
    const double *gamma_fit;     /* [ns*(ns-1)/2*4]                                  */
    int32_t nteta;               /* number of species with thermal diffusion ratios  */
    const int32_t *teta_sp;      /* [nteta]                                          */
    const double *teta_fit;      /* [nteta*ns*4]                                     */
    int32_t viscosity_model;     /* LSM_VISCOSITY_*                                  */
} lsm_transport;

/*
 * Options of laminarSMOKE's per-cell property refresh (solvers/laminarSMOKE/readOptions.H
 * and properties.H), applied on top of the fits above:
 *
 *   diffusivity_model   LSM_DIFFUSIVITY_* (readOptions.H:161-181).  Lewis numbers:
 *                       D_k = lambda / (p psi) / cp / Le_k with psi = cTot MW / p
 *                       (properties.H:153-164); lewis[ns] holds Le_k.
 *   simplified          simplifiedTransportProperties (readOptions.H:129,
 *                       properties.H:130-139): mu = 1.8405e-5 (T / 300)^0.6759 kg/m/s,
 *                       lambda = mu cp / 0.7, before the diffusion coefficients.
 *   soret               soretEffect (properties.H:180-192): the thermal-diffusion output
 *                       holds Dsoret_k = D_k theta_k MW_k / MW [m2/s] for the species with
 *                       thermal diffusion ratios (0 for the others) instead of theta_k.
 *   soot_reference      physicalSootDiffusivity (readOptions.H:283-360, properties.H:166-177):
 *   n_soot, soot_sp,    -1 for none; else D_{soot_sp[i]} = D_{soot_reference} soot_corr[i],
 *   soot_corr           with D_{soot_reference} taken before any correction.
 */
enum {
    LSM_DIFFUSIVITY_MULTICOMPONENT = 0,  /* mixture-averaged (MassDiffusionCoefficients) */
    LSM_DIFFUSIVITY_LEWIS_NUMBERS = 1    /* fixed Lewis numbers                          */
};

typedef struct lsm_transport_options {
    int32_t diffusivity_model;   /* LSM_DIFFUSIVITY_*                               */
    const double *lewis;         /* [ns] Lewis numbers (LSM_DIFFUSIVITY_LEWIS_NUMBERS) */
    int32_t simplified;          /* 0/1                                             */
    int32_t soret;               /* 0/1                                             */
    int32_t soot_reference;      /* species index, or -1                            */
    int32_t n_soot;              /* number of corrected species                     */
    const int32_t *soot_sp;      /* [n_soot] distinct species indices               */
    const double *soot_corr;     /* [n_soot] correction factors                     */
} lsm_transport_options;

#ifdef __cplusplus
}
#endif

#endif /* LSM_TRANSPORT_H */
