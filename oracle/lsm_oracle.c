/*
 * lsm_oracle.c - CPU restatement of the laminarSMOKE reaction step.
 *
 * TEST INFRASTRUCTURE ONLY: this file is the correctness oracle and the CPU
 * baseline.  It is not part of the product; only tests/, the smoke check and the
 * benchmark's CPU-baseline leg may call it.
 *
 * Every function cites the reference lines it follows (OS/ =
 * solvers/openSMOKEpp4laminarSMOKE/, LS/ = solvers/laminarSMOKE/).  Arithmetic
 * is FP64 with the reference's operation order where the reference fixes it;
 * build with -ffp-contract=off so no FMA contraction changes rounding.
 */
#include "lsm_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* OS/math/OpenSMOKEFunctions.hpp:39-40,60-85: MachEpsFloat() returns 1.192092896e-07F,
 * MachEps() returns 2^-52.  OS/math/OpenSMOKEStdInclude.h:53-54. */
#define EPS_D 2.220446049250313e-16
static const double EPS_F = (double)1.192092896e-07F;
static const double TINY_F = (double)1.175494351e-38F;
#define BIG_D 1.7976931348623154e+308
#define MAX_ORDER 5
#define MIN_ORDER 1
#define NZ (MAX_ORDER + 3)

/* ------------------------------------------------------------------------- */
/* Mechanism preprocessing (what the maps build when reading kinetics.xml)   */
/* ------------------------------------------------------------------------- */
typedef struct {
    const lsm_mechanism *m;
    int ns, nr;
    /* NASA coefficient blocks as built by SetCoefficients (hpp:263-343) */
    double *cpHT, *cpLT, *dhHT, *dhLT, *dsHT, *dsLT, *TM, *MW;
    double *dn;          /* change of moles (products - reactants)          */
    int *tb_of, *fo_of;  /* reaction -> third-body / fall-off slot, or -1    */
    /* species-major formation lists, in the reference's class order          */
    int *fptr, *frx;     /* [ns+1], entries                                  */
    double *fnu;         /* signed coefficient (negative for reactants)      */
    int *fcls;           /* 1,2,3,4(=0.5),5(other)                           */
    int rev_sum;         /* test option: formation-rate sums in reverse order */
    /* per-reaction lists sorted by species (reactant entry first on ties):
     * rx_* for dH/dS (signed nu), fw_* reactants, bw_* products               */
    int *rx_ptr, *rx_sp, *fw_ptr, *fw_sp, *bw_ptr, *bw_sp;
    double *rx_nu, *fw_nu, *bw_nu;
    /* forward reaction orders (the reactant list, or the explicit FORD orders) and
     * the non-elementary flags of each side (an order below 1,
     * StoichiometricMap.hpp:256-300) */
    int *od_ptr, *od_sp;
    double *od_nu;
    char *sm_f, *sm_b;
} mech_t;

static int nu_class(double nu) {
    if (nu == 1.0) return 1;
    if (nu == 2.0) return 2;
    if (nu == 3.0) return 3;
    if (nu == 0.5) return 4;
    return 5;
}

static void mech_free(mech_t *M) {
    free(M->cpHT); free(M->cpLT); free(M->dhHT); free(M->dhLT); free(M->dsHT); free(M->dsLT);
    free(M->TM); free(M->MW); free(M->dn); free(M->tb_of); free(M->fo_of);
    free(M->fptr); free(M->frx); free(M->fnu); free(M->fcls);
    free(M->rx_ptr); free(M->rx_sp); free(M->rx_nu); free(M->fw_ptr); free(M->fw_sp);
    free(M->fw_nu); free(M->bw_ptr); free(M->bw_sp); free(M->bw_nu);
    free(M->od_ptr); free(M->od_sp); free(M->od_nu); free(M->sm_f); free(M->sm_b);
}

/* stable sort of (sp, nu) pairs by species */
static void sort_pairs(int n, int *sp, double *nu) {
    for (int a = 1; a < n; a++) {
        int ks = sp[a]; double kn = nu[a]; int b = a - 1;
        while (b >= 0 && sp[b] > ks) { sp[b + 1] = sp[b]; nu[b + 1] = nu[b]; b--; }
        sp[b + 1] = ks; nu[b + 1] = kn;
    }
}

static void mech_init(mech_t *M, const lsm_mechanism *m) {
    const int ns = m->ns, nr = m->nr;
    M->m = m; M->ns = ns; M->nr = nr; M->rev_sum = 0;
    M->cpHT = malloc(sizeof(double) * 5 * ns); M->cpLT = malloc(sizeof(double) * 5 * ns);
    M->dhHT = malloc(sizeof(double) * 6 * ns); M->dhLT = malloc(sizeof(double) * 6 * ns);
    M->dsHT = malloc(sizeof(double) * 6 * ns); M->dsLT = malloc(sizeof(double) * 6 * ns);
    M->TM = malloc(sizeof(double) * ns); M->MW = malloc(sizeof(double) * ns);
    const double one_third = 1. / 3.;
    for (int k = 0; k < ns; k++) {           /* ThermodynamicsMap_CHEMKIN.hpp:263-343 */
        const double *c = m->nasa + LSM_NASA_STRIDE * k;
        for (int i = 0; i < 5; i++) { M->cpHT[5 * k + i] = c[i]; M->cpLT[5 * k + i] = c[7 + i]; }
        const double *H = M->cpHT + 5 * k, *L = M->cpLT + 5 * k;
        double *dh = M->dhHT + 6 * k, *dl = M->dhLT + 6 * k;
        dh[0] = H[0]; dh[1] = 0.50 * H[1]; dh[2] = one_third * H[2]; dh[3] = 0.25 * H[3];
        dh[4] = 0.20 * H[4]; dh[5] = c[5];
        dl[0] = L[0]; dl[1] = 0.50 * L[1]; dl[2] = one_third * L[2]; dl[3] = 0.25 * L[3];
        dl[4] = 0.20 * L[4]; dl[5] = c[12];
        double *sh = M->dsHT + 6 * k, *sl = M->dsLT + 6 * k;
        sh[0] = H[0]; sh[1] = H[1]; sh[2] = 0.50 * H[2]; sh[3] = one_third * H[3];
        sh[4] = 0.25 * H[4]; sh[5] = c[6];
        sl[0] = L[0]; sl[1] = L[1]; sl[2] = 0.50 * L[2]; sl[3] = one_third * L[3];
        sl[4] = 0.25 * L[4]; sl[5] = c[13];
        M->TM[k] = c[16]; M->MW[k] = c[17];
    }
    M->dn = malloc(sizeof(double) * nr);
    for (int j = 0; j < nr; j++) {
        double d = 0.;
        for (int p = m->prod_ptr[j]; p < m->prod_ptr[j + 1]; p++) d += m->prod_nu[p];
        for (int p = m->reac_ptr[j]; p < m->reac_ptr[j + 1]; p++) d -= m->reac_nu[p];
        M->dn[j] = d;
    }
    M->tb_of = malloc(sizeof(int) * nr); M->fo_of = malloc(sizeof(int) * nr);
    for (int j = 0; j < nr; j++) M->tb_of[j] = M->fo_of[j] = -1;
    for (int s = 0; s < m->ntb; s++) M->tb_of[m->tb_rxn[s]] = s;
    for (int s = 0; s < m->nfo; s++) M->fo_of[m->fo_rxn[s]] = s;
    /* Species-major lists (StoichiometricMap.hpp:656-730): for each species the
     * reactant entries by class 1,2,3,0.5,other, then the product entries. */
    const int nnz = m->reac_ptr[nr] + m->prod_ptr[nr];
    M->fptr = calloc(ns + 1, sizeof(int));
    M->frx = malloc(sizeof(int) * (nnz ? nnz : 1));
    M->fnu = malloc(sizeof(double) * (nnz ? nnz : 1));
    M->fcls = malloc(sizeof(int) * (nnz ? nnz : 1));
    int e = 0;
    for (int i = 0; i < ns; i++) {
        M->fptr[i] = e;
        for (int side = 0; side < 2; side++) {
            const int32_t *ptr = side ? m->prod_ptr : m->reac_ptr;
            const int32_t *sp = side ? m->prod_sp : m->reac_sp;
            const double *nu = side ? m->prod_nu : m->reac_nu;
            for (int cls = 1; cls <= 5; cls++)
                for (int j = 0; j < nr; j++)
                    for (int p = ptr[j]; p < ptr[j + 1]; p++)
                        if (sp[p] == i && nu_class(nu[p]) == cls) {
                            M->frx[e] = j; M->fnu[e] = side ? nu[p] : -nu[p]; M->fcls[e] = cls; e++;
                        }
        }
    }
    M->fptr[ns] = e;
    const int nre = m->reac_ptr[nr], npe = m->prod_ptr[nr];
    M->fw_ptr = malloc(sizeof(int) * (nr + 1)); M->bw_ptr = malloc(sizeof(int) * (nr + 1));
    M->rx_ptr = malloc(sizeof(int) * (nr + 1));
    M->fw_sp = malloc(sizeof(int) * (nre + 1)); M->fw_nu = malloc(sizeof(double) * (nre + 1));
    M->bw_sp = malloc(sizeof(int) * (npe + 1)); M->bw_nu = malloc(sizeof(double) * (npe + 1));
    M->rx_sp = malloc(sizeof(int) * (nre + npe + 1)); M->rx_nu = malloc(sizeof(double) * (nre + npe + 1));
    int ef = 0, eb = 0, ex = 0;
    for (int j = 0; j < nr; j++) {
        M->fw_ptr[j] = ef; M->bw_ptr[j] = eb; M->rx_ptr[j] = ex;
        for (int p = m->reac_ptr[j]; p < m->reac_ptr[j + 1]; p++) { M->fw_sp[ef] = m->reac_sp[p]; M->fw_nu[ef++] = m->reac_nu[p]; }
        for (int p = m->prod_ptr[j]; p < m->prod_ptr[j + 1]; p++) { M->bw_sp[eb] = m->prod_sp[p]; M->bw_nu[eb++] = m->prod_nu[p]; }
        sort_pairs(ef - M->fw_ptr[j], M->fw_sp + M->fw_ptr[j], M->fw_nu + M->fw_ptr[j]);
        sort_pairs(eb - M->bw_ptr[j], M->bw_sp + M->bw_ptr[j], M->bw_nu + M->bw_ptr[j]);
        /* merged list: reactants (negative nu) before products for the same species */
        for (int p = M->fw_ptr[j]; p < ef; p++) { M->rx_sp[ex] = M->fw_sp[p]; M->rx_nu[ex++] = -M->fw_nu[p]; }
        for (int p = M->bw_ptr[j]; p < eb; p++) { M->rx_sp[ex] = M->bw_sp[p]; M->rx_nu[ex++] = M->bw_nu[p]; }
        sort_pairs(ex - M->rx_ptr[j], M->rx_sp + M->rx_ptr[j], M->rx_nu + M->rx_ptr[j]);
    }
    M->fw_ptr[nr] = ef; M->bw_ptr[nr] = eb; M->rx_ptr[nr] = ex;
    const int nod = m->ord_ptr ? m->ord_ptr[nr] : 0;
    M->od_ptr = malloc(sizeof(int) * (nr + 1));
    M->od_sp = malloc(sizeof(int) * (nre + nod + 1)); M->od_nu = malloc(sizeof(double) * (nre + nod + 1));
    M->sm_f = calloc(nr + 1, 1); M->sm_b = calloc(nr + 1, 1);
    int eo = 0;
    for (int j = 0; j < nr; j++) {
        M->od_ptr[j] = eo;
        if (m->ord_ptr && m->ord_ptr[j + 1] > m->ord_ptr[j]) {
            for (int p = m->ord_ptr[j]; p < m->ord_ptr[j + 1]; p++) { M->od_sp[eo] = m->ord_sp[p]; M->od_nu[eo++] = m->ord_val[p]; }
        } else {
            for (int p = M->fw_ptr[j]; p < M->fw_ptr[j + 1]; p++) { M->od_sp[eo] = M->fw_sp[p]; M->od_nu[eo++] = M->fw_nu[p]; }
        }
        sort_pairs(eo - M->od_ptr[j], M->od_sp + M->od_ptr[j], M->od_nu + M->od_ptr[j]);
        for (int p = M->od_ptr[j]; p < eo; p++) if (M->od_nu[p] < 1.) M->sm_f[j] = 1;
        for (int p = M->bw_ptr[j]; p < M->bw_ptr[j + 1]; p++) if (M->bw_nu[p] < 1.) M->sm_b[j] = 1;
    }
    M->od_ptr[nr] = eo;
}

/* ------------------------------------------------------------------------- */
/* Thermodynamics: ThermodynamicsMap_CHEMKIN.hpp:436-512                     */
/* ------------------------------------------------------------------------- */
static void thermo_eval(const mech_t *M, double T, double *cp, double *h, double *s) {
    const double T2 = T * T, T3 = T2 * T, T4 = T3 * T, uT = 1. / T, logT = log(T);
    for (int k = 0; k < M->ns; k++) {
        const int hi = T > M->TM[k];
        const double *C = (hi ? M->cpHT : M->cpLT) + 5 * k;
        const double *D = (hi ? M->dhHT : M->dhLT) + 6 * k;
        const double *S = (hi ? M->dsHT : M->dsLT) + 6 * k;
        if (cp) cp[k] = C[0] + T * C[1] + T2 * C[2] + T3 * C[3] + T4 * C[4];
        if (h) h[k] = D[0] + T * D[1] + T2 * D[2] + T3 * D[3] + T4 * D[4] + uT * D[5];
        if (s) s[k] = S[0] * logT + T * S[1] + T2 * S[2] + T3 * S[3] + T4 * S[4] + S[5];
    }
}

void lsm_oracle_species_thermo(const lsm_mechanism *m, double T, double *cp_R, double *h_RT,
                               double *s_R) {
    mech_t M;
    mech_init(&M, m);
    thermo_eval(&M, T, cp_R, h_RT, s_R);
    mech_free(&M);
}

/* ------------------------------------------------------------------------- */
/* Kinetics: KineticsMap_CHEMKIN.hpp                                          */
/* ------------------------------------------------------------------------- */
typedef struct {
    double *h, *s, *cp, *kf, *uKeq, *kmod, *finf, *logFc, *pf, *pb, *r, *R, *dH, *dS;
    double *x, *c;
} kin_ws_t;

static void kin_ws_alloc(kin_ws_t *w, int ns, int nr) {
    w->h = malloc(sizeof(double) * ns); w->s = malloc(sizeof(double) * ns);
    w->cp = malloc(sizeof(double) * ns); w->x = malloc(sizeof(double) * ns);
    w->c = malloc(sizeof(double) * ns); w->R = malloc(sizeof(double) * ns);
    w->kf = malloc(sizeof(double) * nr); w->uKeq = malloc(sizeof(double) * nr);
    w->kmod = malloc(sizeof(double) * nr); w->finf = malloc(sizeof(double) * (nr + 1));
    w->logFc = malloc(sizeof(double) * (nr + 1)); w->pf = malloc(sizeof(double) * nr);
    w->pb = malloc(sizeof(double) * nr); w->r = malloc(sizeof(double) * nr);
    w->dH = malloc(sizeof(double) * nr); w->dS = malloc(sizeof(double) * nr);
}

static void kin_ws_free(kin_ws_t *w) {
    free(w->h); free(w->s); free(w->cp); free(w->x); free(w->c); free(w->R); free(w->kf);
    free(w->uKeq); free(w->kmod); free(w->finf); free(w->logFc); free(w->pf); free(w->pb);
    free(w->r); free(w->dH); free(w->dS);
}

static double cls_val(int cls, double nu, double v) {
    switch (cls) {
    case 1: return v;
    case 2: return v + v;
    case 3: return v + v + v;
    case 4: return 0.5 * v;
    default: return nu * v;
    }
}

/* c^nu by class (StoichiometricMap.hpp:537-552) */
static double conc_pow(double c1, double nu) {
    switch (nu_class(nu)) {
    case 1: return c1;
    case 2: return c1 * c1;
    case 3: return (c1 * c1) * c1;
    case 4: return sqrt(c1);
    default: return pow(c1, nu);
    }
}

/* One factor of the product of concentrations of a non-elementary side:
 * StoichiometricMap::ProductOfConcentrationsForNonElementaryReactions
 * (StoichiometricMap.hpp:617-654): c^lambda for lambda >= 1, below 1 a blend of
 * (c + m/delta)^lambda and a linear branch through 0 around c* = 1e-8 kmol/m3. */
static double order_pow(double C, double lambda) {
    if (lambda >= 1.) return pow(C, lambda);
    const double Cstar = 1.e-8, ALFA = 1.e-5;
    const double H = 1.50 * log(ALFA / (1. - ALFA));
    const double K = 2.00 * log((1. - ALFA) / ALFA) / Cstar;
    const double delta = 1.e9;
    const double mm = (tanh(K * C + H) + 1.) / 2.;
    return mm * pow(C + mm / delta, lambda) + (1. - mm) * pow(Cstar, lambda - 1.) * C;
}

/* h, s must already hold h/RT and s/R at T. */
/* PressureLogarithmicRateExpression::KineticConstant
 * (kernel/kinetics/PressureLogarithmicRateExpression.hpp:104-119): the Arrhenius set
 * of the first/last level outside the pressure range, else ln k interpolated
 * linearly in ln P between the two levels around P. */
static double plog_k(const lsm_mechanism *m, int k, double T, double P) {
    const double *L = m->pl_par + 4 * m->pl_ptr[k];
    const int n = m->pl_ptr[k + 1] - m->pl_ptr[k];
    if (P <= L[0]) return exp(L[1] + L[2] * log(T) - L[3] / T);
    if (P >= L[4 * (n - 1)]) return exp(L[4 * (n - 1) + 1] + L[4 * (n - 1) + 2] * log(T) - L[4 * (n - 1) + 3] / T);
    int i;
    for (i = 0; i < n - 1; i++)
        if (P < L[4 * (i + 1)]) break;
    const double *a = L + 4 * i, *b = L + 4 * (i + 1);
    const double lnkA = a[1] + a[2] * log(T) - a[3] / T;
    const double lnkB = b[1] + b[2] * log(T) - b[3] / T;
    const double lnpA = log(a[0]), lnpB = log(b[0]);
    return exp(lnkA + (lnkB - lnkA) * (log(P) - lnpA) / (lnpB - lnpA));
}

/* ChebyshevPolynomialRateExpression::KineticConstant
 * (kernel/kinetics/ChebyshevPolynomialRateExpression.hpp:95-116, Phi at :187-190). */
static double cheb_k(const lsm_mechanism *m, int k, double T, double P) {
    const int N = m->ch_dims[2 * k], Mm = m->ch_dims[2 * k + 1];
    const double *L = m->ch_lim + 5 * k, *a = m->ch_a + m->ch_ptr[k];
    const double Tmin = L[0], Tmax = L[1];
    const double lPmin = log10(L[2]), lPmax = log10(L[3]);
    const double Tt = (2. / T - 1. / Tmin - 1. / Tmax) / (1. / Tmax - 1. / Tmin);
    const double Pt = (2. * log10(P) - lPmin - lPmax) / (lPmax - lPmin);
    double sum = 0.;
    for (int n = 0; n < N; n++) {
        const double pn = cos((double)n * acos(Tt));
        for (int q = 0; q < Mm; q++) sum += a[n * Mm + q] * pn * cos((double)q * acos(Pt));
    }
    return pow(10., sum) * L[4];
}

/* P is the pressure the kinetics map is set to (SetPressure): it only enters the
 * PLOG and Chebyshev constants. */
static void kinetics_eval(const mech_t *M, kin_ws_t *w, double T, double P, const double *c) {
    const lsm_mechanism *m = M->m;
    const int ns = M->ns, nr = M->nr;
    /* SetTemperature, KineticsMap_CHEMKIN.hpp:271-287 */
    const double uT = 1. / T, logT = log(T);
    const double Patm_over_RT = 101325. / LSM_R_J_KMOL / T;
    const double log_Patm_over_RT = log(Patm_over_RT);
    /* ReactionEnthalpyAndEntropy, StoichiometricMap.hpp:916-1001: accumulation
     * walks species in increasing order; a species on both sides contributes its
     * reactant term first. */
    for (int j = 0; j < nr; j++) {
        double dh = 0., ds = 0.;
        for (int p = M->rx_ptr[j]; p < M->rx_ptr[j + 1]; p++) {
            const int sp = M->rx_sp[p];
            const double nu = M->rx_nu[p];
            if (nu < 0.) {
                dh -= cls_val(nu_class(-nu), -nu, w->h[sp]);
                ds -= cls_val(nu_class(-nu), -nu, w->s[sp]);
            } else {
                dh += cls_val(nu_class(nu), nu, w->h[sp]);
                ds += cls_val(nu_class(nu), nu, w->s[sp]);
            }
        }
        w->dH[j] = dh; w->dS[j] = ds;
    }
    /* KineticConstants, hpp:1198-1212 (Arrhenius + negative A) */
    for (int j = 0; j < nr; j++) {
        w->kf[j] = exp(m->lnA[j] + m->beta[j] * logT - m->E_over_R[j] * uT);
        if (m->rflags[j] & LSM_RXN_NEGATIVE_A) w->kf[j] *= -1.;
    }
    /* hpp:1333-1349: Chebyshev and PLOG constants replace the Arrhenius ones */
    for (int k = 0; k < m->nch; k++) w->kf[m->ch_rxn[k]] = cheb_k(m, k, T, P);
    for (int k = 0; k < m->npl; k++) w->kf[m->pl_rxn[k]] = plog_k(m, k, T, P);
    /* hpp:1215-1222: 1/Kc = exp(-dS/R + dH/RT - dn ln(Patm/RT)); explicitly
     * reversible reactions scale the reverse product by k_rev/k instead
     * (hpp:1242-1253, 1744-1750) */
    for (int j = 0; j < nr; j++) {
        if (!(m->rflags[j] & LSM_RXN_REVERSIBLE)) { w->uKeq[j] = 0.; continue; }
        if (m->rflags[j] & LSM_RXN_EXPLICIT_REV) {
            const double *p = m->rev_arr + 3 * j;
            w->uKeq[j] = exp(p[0] + p[1] * logT - p[2] * uT) / w->kf[j];
        } else {
            w->uKeq[j] = exp(-w->dS[j] + w->dH[j] - log_Patm_over_RT * M->dn[j]);
        }
    }
    /* hpp:1256-1290 (fall-off) and 1292-1326 (CABR): k_inf and Troe log10(Fcent) */
    for (int k = 0; k < m->nfo; k++) {
        const double *p = m->fo_inf + 3 * k;
        const int form = m->fo_type[k] & 3;
        w->finf[k] = exp(p[0] + p[1] * logT - p[2] * uT);
        if (form == LSM_FALLOFF_TROE) {
            const double *t = m->fo_troe + 4 * k;
            double F = (t[3] != 0.) ? (1. - t[0]) * exp(-T / t[1]) + t[0] * exp(-T / t[2]) + exp(-t[3] / T)
                                    : (1. - t[0]) * exp(-T / t[1]) + t[0] * exp(-T / t[2]);
            w->logFc[k] = (F < 1.e-300) ? -300. : log10(F);
        } else if (form == LSM_FALLOFF_SRI) {
            /* hpp:1283-1287: a exp(-b/T) + exp(-T/c), kept in the logFcent slot */
            const double *t = m->fo_sri + 5 * k;
            w->logFc[k] = t[0] * exp(-t[1] / T) + exp(-T / t[2]);
        }
    }
    /* ReactionRates(c), hpp:1683-1690: cTot recomputed as sum(c) */
    double cTot = 0.;
    for (int i = 0; i < ns; i++) cTot += c[i];
    for (int j = 0; j < nr; j++) w->kmod[j] = w->kf[j];
    /* ThirdBodyReactions hpp:1388-1398 and hpp:1703-1708 */
    for (int s = 0; s < m->ntb; s++) {
        double Meff = cTot;
        for (int q = m->tb_ptr[s]; q < m->tb_ptr[s + 1]; q++) Meff += c[m->tb_sp[q]] * m->tb_eff_m1[q];
        w->kmod[m->tb_rxn[s]] *= Meff;
    }
    /* FallOffReactions hpp:1496-1550 and hpp:1710-1716; chemically activated
     * reactions: ChemicallyActivatedBimolecularReactions hpp:1613-1659, 1721-1726 */
    for (int k = 0; k < m->nfo; k++) {
        const int form = m->fo_type[k] & 3, cabr = (m->fo_type[k] & LSM_FALLOFF_CHEMACT) != 0;
        double Mc;
        if (m->fo_single[k] < 0) {
            Mc = cTot;
            for (int q = m->fo_ptr[k]; q < m->fo_ptr[k + 1]; q++) Mc += c[m->fo_sp[q]] * m->fo_eff_m1[q];
        } else {
            Mc = c[m->fo_single[k]] + 1.e-16;
        }
        const int j = m->fo_rxn[k];
        const double Pr = w->kf[j] * Mc / w->finf[k];
        double wF = 1.;
        if (form == LSM_FALLOFF_TROE) {
            const double lf = w->logFc[k];
            if (Pr > 1.e-32 || cabr) {
                const double nT = 0.75 - 1.27 * lf;
                const double cT = -0.4 - 0.67 * lf;
                const double sT = log10(Pr) + cT;
                const double q = sT / (nT - 0.14 * sT);
                wF = pow(10., lf / (1. + q * q));
            } else {
                const double q = 1. / 0.14;
                wF = pow(10., lf / (1. + q * q));
            }
        } else if (form == LSM_FALLOFF_SRI) {   /* hpp:1538-1544, 1645-1651 */
            const double *t = m->fo_sri + 5 * k;
            const double lp = log10(Pr);
            const double xSRI = 1. / (1. + lp * lp);
            wF = pow(w->logFc[k], xSRI) * t[3];
            if (t[4] != 0.) wF *= pow(T, t[4]);
        }
        if (cabr) w->kmod[j] *= (1. / (1. + Pr)) * wF;
        else w->kmod[j] *= w->finf[k] * (Pr / (1. + Pr)) * wF / w->kf[j];
    }
    /* ProductOfConcentrations StoichiometricMap.hpp:512-598 (species order) */
    for (int j = 0; j < nr; j++) {
        double pf = 1., pb = 1.;
        for (int p = M->od_ptr[j]; p < M->od_ptr[j + 1]; p++)
            pf *= M->sm_f[j] ? order_pow(c[M->od_sp[p]], M->od_nu[p]) : conc_pow(c[M->od_sp[p]], M->od_nu[p]);
        if (m->rflags[j] & LSM_RXN_REVERSIBLE)
            for (int p = M->bw_ptr[j]; p < M->bw_ptr[j + 1]; p++)
                pb *= M->sm_b[j] ? order_pow(c[M->bw_sp[p]], M->bw_nu[p]) : conc_pow(c[M->bw_sp[p]], M->bw_nu[p]);
        w->pf[j] = pf; w->pb[j] = pb;
    }
    /* hpp:1737-1767 */
    for (int j = 0; j < nr; j++) {
        double net = w->pf[j];
        if (m->rflags[j] & LSM_RXN_REVERSIBLE) net -= w->pb[j] * w->uKeq[j];
        w->r[j] = net * w->kmod[j];
    }
    /* FormationRatesFromReactionRates StoichiometricMap.hpp:656-730 */
    for (int i = 0; i < ns; i++) {
        double rate = 0.;
        const int e0 = M->fptr[i], e1 = M->fptr[i + 1];
        for (int q = e0; q < e1; q++) {
            const int e = M->rev_sum ? e0 + e1 - 1 - q : q;
            const double nu = M->fnu[e];
            if (nu < 0) rate -= cls_val(M->fcls[e], -nu, w->r[M->frx[e]]);
            else rate += cls_val(M->fcls[e], nu, w->r[M->frx[e]]);
        }
        w->R[i] = rate;
    }
}

void lsm_oracle_kinetics(const lsm_mechanism *m, double T, const double *c, double *kf,
                         double *uKeq, double *kmod, double *r, double *R) {
    lsm_oracle_kinetics_p(m, T, LSM_P_ATM, c, kf, uKeq, kmod, r, R);
}

void lsm_oracle_kinetics_p(const lsm_mechanism *m, double T, double P, const double *c, double *kf,
                           double *uKeq, double *kmod, double *r, double *R) {
    mech_t M;
    kin_ws_t w;
    mech_init(&M, m);
    kin_ws_alloc(&w, m->ns, m->nr);
    thermo_eval(&M, T, w.cp, w.h, w.s);
    kinetics_eval(&M, &w, T, P, c);
    if (kf) memcpy(kf, w.kf, sizeof(double) * m->nr);
    if (uKeq) memcpy(uKeq, w.uKeq, sizeof(double) * m->nr);
    if (kmod) memcpy(kmod, w.kmod, sizeof(double) * m->nr);
    if (r) memcpy(r, w.r, sizeof(double) * m->nr);
    if (R) memcpy(R, w.R, sizeof(double) * m->ns);
    kin_ws_free(&w);
    mech_free(&M);
}

/* ------------------------------------------------------------------------- */
/* Reactor RHS: BatchReactorHomogeneousConstantPressure.H:175-247            */
/* ------------------------------------------------------------------------- */
typedef struct {
    const mech_t *M;
    kin_ws_t w;
    double P;
    int energy;
    int cv;         /* constant-volume reactor */
    double rho0;    /* its density (SetReactor, BatchReactorHomogeneousConstantVolume.H:118-123) */
    double QR;      /* heat release of the last RHS call (reactor QR_, :216-221) */
    long nrhs;
} reactor_t;

static void reactor_rhs(void *vctx, const double *y, double *dy) {
    reactor_t *R = (reactor_t *)vctx;
    const mech_t *M = R->M;
    const int ns = M->ns;
    kin_ws_t *w = &R->w;
    R->nrhs++;
    const double T = y[ns];
    /* MoleFractions_From_MassFractions hpp:429-434, UDot OpenSMOKEUtilities.hpp:339-345 */
    double u = 0.;
    for (int i = 0; i < ns; i++) u += y[i] / M->MW[i];
    const double MW = 1. / u;
    for (int i = 0; i < ns; i++) w->x[i] = (MW * y[i]) / M->MW[i];
    /* constant pressure: cTot = P/(R T); constant volume: cTot = rho0/MW
     * (BatchReactorHomogeneousConstantVolume.H:167-170) */
    const double cTot = R->cv ? R->rho0 / MW : R->P / LSM_R_J_KMOL / T;
    for (int i = 0; i < ns; i++) w->c[i] = cTot * w->x[i];
    const double rho = R->cv ? R->rho0 : cTot * MW;
    thermo_eval(M, T, w->cp, w->h, w->s);
    /* the map's pressure: P0, or cTot R T in the constant-volume reactor
     * (BatchReactorHomogeneousConstantVolume.H:166-176) */
    kinetics_eval(M, w, T, R->cv ? cTot * LSM_R_J_KMOL * T : R->P, w->c);
    for (int i = 0; i < ns; i++) dy[i] = M->MW[i] * w->R[i] / rho;
    dy[ns] = 0.;
    if (R->energy && R->cv) {
        /* BatchReactorHomogeneousConstantVolume.H:185-196:
         * dT/dt = (QR + R T sum_k R_k) / (rho0 cv), cv = (CpMixMolar - R)/MW */
        double cpd = 0.;
        for (int i = 0; i < ns; i++) cpd += w->cp[i] * w->x[i];
        const double CvMixMass = (cpd * LSM_R_J_KMOL - LSM_R_J_KMOL) / MW;
        double hd = 0., sumR = 0.;
        for (int i = 0; i < ns; i++) { hd += w->R[i] * w->h[i]; sumR += w->R[i]; }
        R->QR = -hd * LSM_R_J_KMOL * T;
        dy[ns] = (R->QR + LSM_R_J_KMOL * T * sumR) / (rho * CvMixMass);
    } else if (R->energy) {
        double cpd = 0.;
        for (int i = 0; i < ns; i++) cpd += w->cp[i] * w->x[i];
        const double CpMixMolar = cpd * LSM_R_J_KMOL;
        const double CpMixMass = CpMixMolar / MW;
        double hd = 0.;
        for (int i = 0; i < ns; i++) hd += w->R[i] * w->h[i];
        R->QR = -hd * LSM_R_J_KMOL * T;       /* KineticsMap_CHEMKIN.hpp:1961-1964 */
        dy[ns] = R->QR / (rho * CpMixMass);
    }
}

static double rhs_once(const lsm_mechanism *m, double P, int cv, double rho0, int energy, const double *y,
                       double *f) {
    mech_t M;
    reactor_t R;
    mech_init(&M, m);
    R.M = &M; R.P = P; R.energy = energy; R.cv = cv; R.rho0 = rho0; R.QR = 0.; R.nrhs = 0;
    kin_ws_alloc(&R.w, m->ns, m->nr);
    reactor_rhs(&R, y, f);
    kin_ws_free(&R.w);
    mech_free(&M);
    return R.QR;
}

double lsm_oracle_rhs(const lsm_mechanism *m, double P, int energy, const double *y, double *f) {
    return rhs_once(m, P, 0, 0., energy, y, f);
}

double lsm_oracle_rhs_cv(const lsm_mechanism *m, double rho0, int energy, const double *y, double *f) {
    return rhs_once(m, 101325., 1, rho0, energy, y, f);
}

/* ------------------------------------------------------------------------- */
/* Dense LU (KernelDense.hpp:196-246 call Eigen PartialPivLU / FullPivLU)    */
/* ------------------------------------------------------------------------- */
int lsm_oracle_lu_factor(int n, double *A, int32_t *rp, int32_t *cp, int full) {
    int first_zero = 0;
    for (int i = 0; i < n; i++) { rp[i] = i; if (cp) cp[i] = i; }
    for (int k = 0; k < n; k++) {
        int p = k, q = k;
        double best = -1.;
        if (!full) {
            for (int i = k; i < n; i++) {
                const double a = fabs(A[i * n + k]);
                if (a > best) { best = a; p = i; }
            }
        } else {
            /* column-major scan of the trailing block: first maximum wins */
            for (int j = k; j < n; j++)
                for (int i = k; i < n; i++) {
                    const double a = fabs(A[i * n + j]);
                    if (a > best) { best = a; p = i; q = j; }
                }
        }
        if (p != k) {
            for (int j = 0; j < n; j++) { double t = A[k * n + j]; A[k * n + j] = A[p * n + j]; A[p * n + j] = t; }
            int t = rp[k]; rp[k] = rp[p]; rp[p] = t;
        }
        if (full && q != k) {
            for (int i = 0; i < n; i++) { double t = A[i * n + k]; A[i * n + k] = A[i * n + q]; A[i * n + q] = t; }
            int t = cp[k]; cp[k] = cp[q]; cp[q] = t;
        }
        const double piv = A[k * n + k];
        if (piv == 0.) { if (!first_zero) first_zero = k + 1; continue; }
        for (int i = k + 1; i < n; i++) {
            const double l = A[i * n + k] / piv;
            A[i * n + k] = l;
            for (int j = k + 1; j < n; j++) A[i * n + j] -= l * A[k * n + j];
        }
    }
    return first_zero;
}

void lsm_oracle_lu_solve(int n, const double *LU, const int32_t *rp, const int32_t *cp, double *b,
                         int full) {
    double *t = malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) t[i] = b[rp[i]];
    for (int i = 0; i < n; i++)
        for (int k = 0; k < i; k++) t[i] -= LU[i * n + k] * t[k];
    for (int i = n - 1; i >= 0; i--) {
        for (int k = i + 1; k < n; k++) t[i] -= LU[i * n + k] * t[k];
        t[i] /= LU[i * n + i];
    }
    if (full && cp) for (int i = 0; i < n; i++) b[cp[i]] = t[i];
    else memcpy(b, t, sizeof(double) * n);
    free(t);
}

/* ------------------------------------------------------------------------- */
/* Gear/BDF: MultiValueSolver.hpp + MethodGear.hpp + KernelDense.hpp          */
/* ------------------------------------------------------------------------- */
enum { H_DEC = 0, H_CONST = 1, H_INC = 2 };                 /* OdeSolverUtilities.h:43-48 */
enum { O_DEC = 0, O_CONST = 1, O_INC = 2 };
enum { J_CHANGE = 0, J_MODIFIED = 1, J_OK = 2 };
enum { M_TO_FACTOR = 0, M_FACTORED = 1 };
enum { CONV_FAIL = 0, CONV_OK = 1 };
enum { ST_INIT = 1, ST_CONT = 2, ST_MAXSTEPS = -1, ST_STRICT = -2, ST_ERRFAIL = -4,
       ST_CONVFAIL = -5, ST_SMALLH = -6, ST_ILLEGAL = -9 };

typedef struct {
    int ne;
    lsm_oracle_rhs_fn rhs;
    void *ctx;
    /* options */
    double abs_tol, rel_tol;
    int full_pivoting, min_c, max_c;
    const double *ymin, *ymax;
    unsigned max_steps;
    /* vectors */
    double *z[NZ], *v[NZ], *b, *y, *va, *deltab, *f, *vb, *ooe, *y0, *aux, *J, *G;
    int32_t *rp, *cp;
    /* method constants (MethodGear.hpp:126-194) */
    double r[MAX_ORDER + 1][MAX_ORDER], Ep[MAX_ORDER + 1], alfa2[MAX_ORDER + 1];
    double deltaAlfa1, deltaAlfa3;
    /* state */
    double t, t0, tOut, tInMeshPoint, tStabilize;
    double h, hNextStep, hNordsieck, hPreviousStep, hScale, hScaleMax, first_step_size;
    double max_step_size, min_step_size, hMinUsed, hMaxUsed;
    int p, orderInNextStep, maximum_order, status;
    int odeHStatus, odeOrderStatus, factorizationStatus, jacobianStatus, convergenceStatus;
    unsigned numberOfSteps, numberOfFunctionCalls, numberOfLinearSystemSolutions, numberOfJacobians;
    unsigned numberOfMatrixFactorizations, numberOfDecreasedSteps, numberOfIncreasedSteps;
    unsigned numberOfConvergenceFailuresForOrderMax, numberOfFunctionCallsForJacobian;
    unsigned iterConvergence, iterConvergenceRate, iterConvergenceFailure, iterErrorFailure, iterOrder;
    unsigned maxIterationsJacobian, maxConvergenceIterations, stepOfLastJacobian, stepOfLastFactorization;
    unsigned numberOfStepsWithoutChanges, numberOfConvergenceFailure, numberOfConvergenceSuccess;
    unsigned numberOfErrorCheckFailure, numberOfErrorCheckSuccess;
    unsigned iterConstSteps, iterMinimumOrder, maxOrderUsed;
    int fixedOrderMin;
    double jac_scale;          /* factor on the FD increments (1: the reference's), test option */
    long clamps, newton_iters;
} bdf_t;

static unsigned factorial(unsigned n) { unsigned f = 1; for (unsigned i = 2; i <= n; i++) f *= i; return f; }

/* ErrorControl, OpenSMOKEUtilities.hpp:689-701 */
static double wrms(int n, const double *w, const double *e) {
    double sum = 0.;
    for (int i = 0; i < n; i++) { const double c = w[i] * e[i]; sum += c * c; }
    return sqrt(sum / (double)n);
}

/* Fresh-solver state: MultiValueSolver constructor + MethodGear::MemoryAllocationMethod
 * (MethodGear.hpp:126-194).  alfa2_ and maximum_order_ start from their defaults for
 * every cell (the reference carries them over between cells; that leak is not
 * reproduced). */
static void bdf_fresh(bdf_t *s) {
    const int ne = s->ne;
    for (int i = 0; i < NZ; i++) { memset(s->z[i], 0, 8 * ne); memset(s->v[i], 0, 8 * ne); }
    s->deltaAlfa1 = 0.05; s->deltaAlfa3 = 0.25;
    for (int i = 0; i <= MAX_ORDER; i++) s->alfa2[i] = 0.8;
    s->maximum_order = MAX_ORDER;     /* MultiValueSolver::SetDefaultValues :64-66 */
    s->max_step_size = BIG_D;
    s->max_steps = 500000;
    s->clamps = 0; s->newton_iters = 0;
    s->jac_scale = 1.;
}

static void bdf_alloc(bdf_t *s, int ne) {
    memset(s, 0, sizeof(*s));
    s->ne = ne;
    for (int i = 0; i < NZ; i++) { s->z[i] = calloc(ne, sizeof(double)); s->v[i] = calloc(ne, sizeof(double)); }
    s->b = calloc(ne, 8); s->y = calloc(ne, 8); s->va = calloc(ne, 8); s->deltab = calloc(ne, 8);
    s->f = calloc(ne, 8); s->vb = calloc(ne, 8); s->ooe = calloc(ne, 8); s->y0 = calloc(ne, 8);
    s->aux = calloc(ne, 8); s->J = calloc((size_t)ne * ne, 8); s->G = calloc((size_t)ne * ne, 8);
    s->rp = calloc(ne, sizeof(int32_t)); s->cp = calloc(ne, sizeof(int32_t));
    for (int j = 0; j <= MAX_ORDER; j++) s->Ep[j] = factorial(j);
    for (int i = 0; i <= MAX_ORDER; i++) for (int j = 0; j < MAX_ORDER; j++) s->r[i][j] = 0.;
    s->r[0][0] = 1.; s->r[1][0] = 1.;
    for (int j = 2; j <= MAX_ORDER; j++) {
        s->r[0][j - 1] = factorial(j);
        for (int i = 2; i <= j; i++) s->r[i - 1][j - 1] = (double)j * s->r[i - 1][j - 2] + s->r[i - 2][j - 2];
        s->r[j][j - 1] = 1.;
    }
    double sum = 1.;
    for (int j = 2; j <= MAX_ORDER; j++) {
        sum += 1. / (double)j;
        for (int i = 1; i <= j + 1; i++) s->r[i - 1][j - 1] /= (factorial(j) * sum);
    }
    bdf_fresh(s);
}

static void bdf_free(bdf_t *s) {
    for (int i = 0; i < NZ; i++) { free(s->z[i]); free(s->v[i]); }
    free(s->b); free(s->y); free(s->va); free(s->deltab); free(s->f); free(s->vb); free(s->ooe);
    free(s->y0); free(s->aux); free(s->J); free(s->G); free(s->rp); free(s->cp);
}

static void eval(bdf_t *s, const double *y, double *f) { s->rhs(s->ctx, y, f); }

/* CheckConstraints MethodGear.hpp:519-536 */
static void check_constraints(bdf_t *s, double *y) {
    if (s->min_c) for (int i = 0; i < s->ne; i++) if (y[i] < s->ymin[i]) { y[i] = s->ymin[i]; s->clamps++; }
    if (s->max_c) for (int i = 0; i < s->ne; i++) if (y[i] > s->ymax[i]) { y[i] = s->ymax[i]; s->clamps++; }
}

/* CheckIllegalConstraints MethodGear.hpp:539-569 */
static int illegal_constraints(bdf_t *s, const double *y) {
    int n = 0;
    if (s->min_c) for (int i = 0; i < s->ne; i++) if (y[i] < s->ymin[i]) n++;
    if (s->max_c) for (int i = 0; i < s->ne; i++) if (y[i] > s->ymax[i]) n++;
    return n;
}

/* CalculatesOneOverEpsilon MultiValueSolver.hpp:620-646 (scalar tolerances) */
static void one_over_epsilon(bdf_t *s, const double *w) {
    for (int i = 0; i < s->ne; i++) s->ooe[i] = s->abs_tol + s->rel_tol * fabs(w[i]);
    for (int i = 0; i < s->ne; i++) s->ooe[i] = 1. / s->ooe[i];
}

/* KernelDense::NumericalJacobian KernelDense.hpp:112-194 */
static void numerical_jacobian(bdf_t *s, double *y, const double *f, double h, const double *e) {
    const int ne = s->ne;
    const double ZERO_DER = sqrt(TINY_F), ETA2 = sqrt(EPS_D), BETA = 1.e+3 * EPS_D;
    double hf = BETA * fabs(h) * wrms(ne, f, e) * (double)ne;
    if (hf < 1.e-10) hf = 1.;
    for (int j = 0; j < ne; j++) {
        const double yh = y[j];
        double hJ;
        if (!s->max_c) {
            hJ = ETA2 * fabs(fmax(yh, 1. / e[j]));
            const double hJf = hf / e[j];
            hJ = fmax(hJ, hJf);
            hJ = fmax(hJ, ZERO_DER);
            hJ = fmin(hJ, 0.001 + 0.001 * fabs(yh));
        } else {
            const double a = fabs(yh), bb = fabs(1. / e[j]);
            hJ = ETA2 * (a > bb ? a : bb);          /* OpenSMOKE::MaxAbs, OpenSMOKEUtilities.hpp:90-93 */
            const double hJf = hf / e[j];
            hJ = fmax(hJ, hJf);
            hJ = fmax(hJ, ZERO_DER);
            hJ *= s->jac_scale;
            if (yh + hJ > s->ymax[j]) hJ = -hJ;
        }
        y[j] += hJ;
        eval(s, y, s->aux);
        y[j] = yh;
        for (int i = 0; i < ne; i++) s->aux[i] -= f[i];
        const double hInv = 1. / hJ;
        for (int i = 0; i < ne; i++) s->J[i * ne + j] = hInv * s->aux[i];
    }
    s->numberOfFunctionCallsForJacobian += ne;
}

/* BuildAndFactorizeMatrixG KernelDense.hpp:196-221 */
static void build_and_factorize(bdf_t *s, double hr0) {
    const int ne = s->ne;
    for (int i = 0; i < ne * ne; i++) s->G[i] = s->J[i] * -hr0;
    for (int i = 0; i < ne; i++) s->G[i * ne + i] += 1.;
    lsm_oracle_lu_factor(ne, s->G, s->rp, s->cp, s->full_pivoting);
}

static void solve_linear(bdf_t *s, double *db) {
    lsm_oracle_lu_solve(s->ne, s->G, s->rp, s->cp, db, s->full_pivoting);
}

/* MethodGear::ResetMethod MethodGear.hpp:80-124 (+ KernelDense::ResetKernel :44-58) */
static void reset_method(bdf_t *s) {
    s->p = 1; s->iterOrder = 0; s->maxOrderUsed = 1; s->orderInNextStep = 1;
    s->maxConvergenceIterations = 4;
    s->maxIterationsJacobian = 20 + 5 * (s->ne - 3);
    if (s->maxIterationsJacobian > 50 && s->ne < 80) s->maxIterationsJacobian = 50;
    s->odeHStatus = H_CONST; s->odeOrderStatus = O_CONST; s->status = ST_INIT;
    s->factorizationStatus = M_TO_FACTOR; s->jacobianStatus = J_CHANGE;
    s->convergenceStatus = CONV_OK;
    s->h = 0.; s->min_step_size = 0.; s->hNextStep = 0.; s->hScale = 0.;
    s->numberOfSteps = s->numberOfFunctionCalls = s->numberOfLinearSystemSolutions = 0;
    s->numberOfJacobians = s->numberOfMatrixFactorizations = 0;
    s->numberOfDecreasedSteps = s->numberOfIncreasedSteps = 0;
    s->numberOfConvergenceFailuresForOrderMax = 0; s->numberOfFunctionCallsForJacobian = 0;
    s->iterConvergence = s->iterConvergenceRate = s->iterConvergenceFailure = 0;
    s->stepOfLastJacobian = 0;
}

/* MultiValueSolver::Reset :107-131 */
static void reset_solver(bdf_t *s) {
    s->hPreviousStep = 0.; s->hNordsieck = 0.; s->hMinUsed = BIG_D; s->hMaxUsed = 0.;
    s->hScaleMax = 1000.; s->orderInNextStep = 1; s->fixedOrderMin = 0;
    s->numberOfStepsWithoutChanges = s->numberOfConvergenceFailure = s->numberOfConvergenceSuccess = 0;
    s->numberOfErrorCheckFailure = s->numberOfErrorCheckSuccess = 0;
    s->iterConstSteps = 0; s->iterMinimumOrder = 0;
}

/* SetInitialConditions MultiValueSolver.hpp:368-407 */
static void set_initial_conditions(bdf_t *s, double t0, const double *y0) {
    s->tOut = t0; s->t = t0; s->t0 = t0; s->tInMeshPoint = t0; s->tStabilize = t0;
    reset_method(s);
    memcpy(s->y0, y0, sizeof(double) * s->ne);
    reset_solver(s);
    eval(s, s->y0, s->f);
}

/* CalculateInitialStepSize MultiValueSolver.hpp:649-685 */
static double initial_step_size(bdf_t *s) {
    double lo, hi;
    if (s->t0 != 0.) { lo = 100. * EPS_D * fabs(s->t0); hi = fabs(s->t0); }
    else { lo = EPS_D; hi = 1.e-4; }
    one_over_epsilon(s, s->y0);
    double h0 = 0.5 / (wrms(s->ne, s->f, s->ooe) + EPS_F);
    if (h0 < lo) h0 = lo;
    if (h0 > hi) h0 = hi;
    h0 *= fabs(s->tOut - s->t0) / (s->tOut - s->t0);
    return h0;
}

/* MethodGear::FindCorrection MethodGear.hpp:196-341 */
static unsigned find_correction(bdf_t *s) {
    const int ne = s->ne;
    const double *e = s->ooe;
    memcpy(s->vb, s->v[0], sizeof(double) * ne);
    check_constraints(s, s->vb);
    eval(s, s->vb, s->f);
    s->numberOfFunctionCalls++;
    if (s->numberOfSteps >= s->stepOfLastJacobian + s->maxIterationsJacobian) s->jacobianStatus = J_CHANGE;
    if (s->jacobianStatus == J_CHANGE) {
        s->stepOfLastJacobian = s->numberOfSteps;
        s->jacobianStatus = J_MODIFIED;
        numerical_jacobian(s, s->vb, s->f, s->h, e);
        s->numberOfJacobians++;
        s->factorizationStatus = M_TO_FACTOR;
    }
    const double r0 = s->r[0][s->p - 1];
    const double hr0 = s->h * r0;
    if (s->factorizationStatus == M_FACTORED && s->iterOrder == 0) s->factorizationStatus = M_TO_FACTOR;
    if (s->factorizationStatus == M_TO_FACTOR) {
        s->stepOfLastFactorization = s->numberOfSteps;
        s->numberOfMatrixFactorizations++;
        build_and_factorize(s, hr0);
        s->factorizationStatus = M_FACTORED;
    }
    for (int i = 0; i < ne; i++) s->deltab[i] = s->f[i] * s->h;
    for (int i = 0; i < ne; i++) s->deltab[i] -= s->v[1][i];
    solve_linear(s, s->deltab);
    s->numberOfLinearSystemSolutions++;
    memcpy(s->b, s->deltab, sizeof(double) * ne);
    double old = 0.;
    for (unsigned k = 1; k <= s->maxConvergenceIterations; k++) {
        const double cc = r0 * wrms(ne, s->deltab, e);
        if (cc < 1.) { s->convergenceStatus = CONV_OK; s->newton_iters += k; return k; }
        if (k > 1 && cc > 2. * old) { s->convergenceStatus = CONV_FAIL; s->newton_iters += k; return k; }
        old = cc;
        for (int i = 0; i < ne; i++) s->va[i] = s->b[i] * r0;
        for (int i = 0; i < ne; i++) s->va[i] += s->v[0][i];
        if (s->min_c)
            for (int j = 0; j < ne; j++)
                if (s->va[j] < s->ymin[j]) { s->va[j] = s->ymin[j]; s->b[j] = (s->ymin[j] - s->z[0][j]) / r0; s->clamps++; }
        if (s->max_c)
            for (int j = 0; j < ne; j++)
                if (s->va[j] > s->ymax[j]) { s->va[j] = s->ymax[j]; s->b[j] = (s->ymax[j] - s->z[0][j]) / r0; s->clamps++; }
        eval(s, s->va, s->f);
        s->numberOfFunctionCalls++;
        for (int i = 0; i < ne; i++) s->deltab[i] = s->f[i] * s->h;
        for (int i = 0; i < ne; i++) s->deltab[i] -= s->v[1][i];
        for (int i = 0; i < ne; i++) s->deltab[i] -= s->b[i];
        solve_linear(s, s->deltab);
        s->numberOfLinearSystemSolutions++;
        for (int i = 0; i < ne; i++) s->b[i] += s->deltab[i];
    }
    s->convergenceStatus = CONV_FAIL;
    s->newton_iters += s->maxConvergenceIterations;
    return s->maxConvergenceIterations;
}

/* MethodGear::WhatToDoInCaseOfConvergenceFailure MethodGear.hpp:344-423 */
static void convergence_failure(bdf_t *s) {
    s->hScale = 0.25;
    if (s->numberOfSteps < 10) s->hScale = 0.1;
    s->hScale = fmax(s->hScale, s->min_step_size / fabs(s->h));
    s->odeHStatus = H_DEC;
    if (s->p > MIN_ORDER) { s->p -= 1; s->odeOrderStatus = O_DEC; }
    s->alfa2[4] -= 0.01;
    s->alfa2[5] -= 0.02;
    if (s->alfa2[4] < 0.) {
        s->alfa2[4] = 0.; s->maximum_order = 3;
        if (s->p > 3) { s->p = 3; s->odeOrderStatus = O_DEC; }
    }
    if (s->alfa2[5] < 0.) {
        s->alfa2[5] = 0.; s->maximum_order = 4;
        if (s->p > 4) { s->p = 4; s->odeOrderStatus = O_DEC; }
    }
    if (s->numberOfSteps > 5) {
        s->numberOfConvergenceFailuresForOrderMax++;
        if (s->numberOfConvergenceFailuresForOrderMax % 5 == 0)
            if (s->p != MIN_ORDER) { s->p = MIN_ORDER; s->odeOrderStatus = O_DEC; }
        if (s->numberOfConvergenceFailuresForOrderMax % 10 == 0 && s->tStabilize != s->tInMeshPoint) {
            eval(s, s->z[0], s->va);
            s->numberOfFunctionCalls++;
            for (int i = 0; i < s->ne; i++) s->z[1][i] = s->va[i] * s->h;
            s->tStabilize = s->tInMeshPoint;
        }
    }
    if (s->numberOfSteps > s->stepOfLastJacobian + 50 / 10) s->jacobianStatus = J_CHANGE;
    if (s->numberOfSteps != s->stepOfLastJacobian && s->iterConvergenceFailure == 50) s->jacobianStatus = J_CHANGE;
    s->hNextStep *= s->hScale;
    s->numberOfDecreasedSteps++;
}

/* MethodGear::ConvergenceRate MethodGear.hpp:426-516 */
static void convergence_rate(bdf_t *s) {
    const int ne = s->ne;
    if (s->iterErrorFailure > 0 || s->iterConvergenceFailure > 0) s->iterConvergenceRate = 0;
    if (s->iterConvergence <= 1) return;
    for (int i = 0; i < ne; i++) s->vb[i] = s->v[0][i] - s->z[0][i];
    for (int i = 0; i < ne; i++) {
        double acc = 0.;
        for (int j = 0; j < ne; j++) acc += s->J[i * ne + j] * s->vb[j];
        s->va[i] = acc;
    }
    for (int i = 0; i < ne; i++) s->va[i] *= s->h;
    for (int i = 0; i < ne; i++) s->va[i] += s->z[1][i];
    for (int i = 0; i < ne; i++) s->va[i] = s->v[1][i] - s->va[i];
    double normd = 0., normf = 0.;
    for (int i = 0; i < ne; i++) normd += fabs(s->va[i]);
    for (int i = 0; i < ne; i++) normf += fabs(s->v[1][i]);
    if (normd > 3. * normf + .1 && s->iterConvergence > 2) {
        if (s->numberOfSteps > s->stepOfLastJacobian + 50 / 10) s->jacobianStatus = J_CHANGE;
        s->hScale = 0.5; s->odeHStatus = H_DEC; s->hNextStep = s->h * s->hScale;
        s->numberOfDecreasedSteps++;
        if (s->p > 1) { s->orderInNextStep = s->p - 1; s->odeOrderStatus = O_DEC; }
        s->iterConvergenceRate = 0; s->iterConvergenceFailure = 1;
        return;
    }
    if (normd > normf + 0.1) {
        s->iterConvergenceRate++;
        if (s->iterConvergenceRate >= 5) {
            s->hScale = 0.8; s->odeHStatus = H_DEC; s->hNextStep = s->h * s->hScale;
            s->numberOfDecreasedSteps++;
            s->iterConvergenceRate = 0; s->iterConvergenceFailure = 1;
        }
        if (s->numberOfSteps > s->stepOfLastJacobian + 50 / 10 && s->iterConvergence > 2) {
            s->jacobianStatus = J_CHANGE;
            s->iterConvergenceRate = 0; s->iterConvergenceFailure = 1;
        }
    }
}

/* MultiValueSolver::MultiValueStep MultiValueSolver.hpp:1213-1328 */
static void multi_value_step(bdf_t *s) {
    const int ne = s->ne, p = s->p;
    for (int i = 0; i <= p; i++) memcpy(s->v[i], s->z[i], sizeof(double) * ne);
    for (int i = 0; i < p; i++)
        for (int j = p - 1; j >= i; j--)
            for (int q = 0; q < ne; q++) s->v[j][q] += s->v[j + 1][q];
    if (s->min_c)
        for (int j = 0; j < ne; j++)
            if (s->v[0][j] < s->ymin[j] && fabs(s->v[0][j] - s->ymin[j]) * s->ooe[j] < (double)ne) {
                s->v[0][j] = s->ymin[j];
                s->v[1][j] = s->ymin[j] - s->z[0][j];
                for (int i = 2; i <= p; i++) s->v[i][j] = 0.;
                s->clamps++;
            }
    if (s->max_c)
        for (int j = 0; j < ne; j++)
            if (s->v[0][j] > s->ymax[j] && fabs(s->v[0][j] - s->ymax[j]) * s->ooe[j] < (double)ne) {
                s->v[0][j] = s->ymax[j];
                s->v[1][j] = s->ymax[j] - s->z[0][j];
                for (int i = 2; i <= p; i++) s->v[i][j] = 0.;
                s->clamps++;
            }
    s->iterConvergence = find_correction(s);
    if (s->convergenceStatus == CONV_FAIL) return;
    const double r0 = s->r[0][p - 1];
    for (int i = 0; i < ne; i++) s->va[i] = s->b[i] * r0;
    if (s->min_c)
        for (int j = 0; j < ne; j++)
            if (s->v[0][j] + s->va[j] < s->ymin[j]) {
                s->b[j] = s->ymin[j] - s->z[0][j] - s->v[1][j];
                s->v[0][j] = s->ymin[j] - s->va[j];
                if (s->v[0][j] < s->ymin[j]) s->v[0][j] = s->ymin[j];
                s->clamps++;
            }
    if (s->max_c)
        for (int j = 0; j < ne; j++)
            if (s->v[0][j] + s->va[j] > s->ymax[j]) {
                s->b[j] = s->ymax[j] - s->z[0][j] - s->v[1][j];
                s->v[0][j] = s->ymax[j] - s->va[j];
                if (s->v[0][j] > s->ymax[j]) s->v[0][j] = s->ymax[j];
                s->clamps++;
            }
    for (int i = 0; i < ne; i++) s->v[0][i] += s->va[i];
    for (int i = 0; i < ne; i++) s->v[1][i] += s->b[i];
    for (int k = 2; k <= p; k++)
        for (int i = 0; i < ne; i++) { s->va[i] = s->b[i] * s->r[k][p - 1]; s->v[k][i] += s->va[i]; }
    for (int i = 0; i < ne; i++) { s->va[i] = s->v[p][i] - s->z[p][i]; s->v[p + 1][i] = s->va[i] / ((double)p + 1.); }
    if (s->iterOrder >= (unsigned)p)
        for (int i = 0; i < ne; i++) { s->va[i] = s->v[p + 1][i] - s->z[p + 1][i]; s->v[p + 2][i] = s->va[i] / ((double)p + 2.); }
}

/* MultiValueSolver::NewOrderNewH MultiValueSolver.hpp:1043-1210 (CPU-time
 * heuristics dropped: MultiValueSolver::cpuTimeToFactorize_ is always 0, :133,
 * so those branches never fire and scl stays 1.2) */
static void new_order_new_h(bdf_t *s) {
    const int ne = s->ne, p = s->p;
    double hm = 0., hc = 0., hp = 0., alfa1 = 0., alfa3 = 0.;
    if (p > MIN_ORDER && s->odeOrderStatus != O_INC) {
        hm = s->Ep[p - 1] * wrms(ne, s->v[p], s->ooe) + EPS_F;
        hm = 1. / hm;
        alfa1 = s->alfa2[p - 1] - s->deltaAlfa1;
        hm = alfa1 * pow(hm, 1. / (double)p);
    }
    hc = s->Ep[p] * wrms(ne, s->v[p + 1], s->ooe) + EPS_F;
    hc = 1. / hc;
    hc = s->alfa2[p] * pow(hc, 1. / (double)(p + 1));
    if (p < s->maximum_order && s->odeOrderStatus != O_DEC) {
        hp = s->Ep[p + 1] * wrms(ne, s->v[p + 2], s->ooe) + EPS_F;
        hp = 1. / hp;
        alfa3 = s->alfa2[p + 1] - s->deltaAlfa3;
        hp = alfa3 * pow(hp, 1. / (double)(p + 2));
    }
    if (hm > hc && hm > hp) {
        s->hScale = (.75 / alfa1) * hm; s->orderInNextStep = p - 1; s->odeOrderStatus = O_DEC;
    } else if (hc > hm && hc > hp) {
        s->hScale = (.75 / s->alfa2[p]) * hc; s->orderInNextStep = p; s->odeOrderStatus = O_CONST;
    } else {
        s->hScale = (.75 / alfa3) * hp; s->orderInNextStep = p + 1; s->odeOrderStatus = O_INC;
    }
    if (s->fixedOrderMin) {
        s->iterMinimumOrder++;
        if (s->iterMinimumOrder < 2000) {
            s->hScale = (.75 / s->alfa2[p]) * hc; s->orderInNextStep = p; s->odeOrderStatus = O_CONST;
        } else s->fixedOrderMin = 0;
    }
    const double scl = 1.2;
    if (s->hScale < scl) {
        s->iterConstSteps++;
        s->hScale = 1.; s->odeHStatus = H_CONST; s->orderInNextStep = p; s->odeOrderStatus = O_CONST;
        s->hNextStep = s->h;
    } else {
        s->iterConstSteps = 0;
        s->hScale = fmin(s->hScale, s->hScaleMax);
        s->odeHStatus = H_INC;
        s->hNextStep = s->h * s->hScale;
        if (s->hNextStep > s->max_step_size) s->hNextStep = s->max_step_size;
    }
    if (s->iterConstSteps > 500 ||
        (s->numberOfSteps % 1000 == 0 && s->numberOfSteps != 0 &&
         (fabs(s->h) < 1.e-7 || fabs(s->h) < 1.e-5 * s->t))) {
        if (p > MIN_ORDER) {
            s->fixedOrderMin = 1; s->iterMinimumOrder = 0;
            s->orderInNextStep = MIN_ORDER; s->odeOrderStatus = O_DEC;
            s->hScale = 0.5; s->hNextStep = s->h * s->hScale; s->odeHStatus = H_DEC;
        }
    }
    if (s->odeHStatus == H_INC) s->numberOfIncreasedSteps++;
    if (s->odeHStatus == H_DEC) s->numberOfDecreasedSteps++;
}

/* MultiValueSolver::DriverStep MultiValueSolver.hpp:754-1040 */
static void driver_step(bdf_t *s) {
    const int ne = s->ne;
    s->iterConvergenceFailure = 0; s->iterErrorFailure = 0; s->convergenceStatus = CONV_OK;
    if (s->odeOrderStatus != O_CONST || s->odeHStatus != H_CONST) s->iterOrder = 0;
    s->p = s->orderInNextStep;
    s->odeOrderStatus = O_CONST;
    s->hScaleMax = 2.;
    if (s->numberOfSteps < 10) s->hScaleMax = 1000.;
    else if (s->numberOfSteps > 10 && s->numberOfSteps < 20) s->hScaleMax = 10.;
    for (;;) {
        if (s->odeHStatus != H_CONST) {
            s->hScale = s->hNextStep / s->hNordsieck;
            double hr = 1.;
            for (int i = 1; i <= s->p; i++) {
                hr *= s->hScale;
                for (int q = 0; q < ne; q++) s->z[i][q] *= hr;
            }
            check_constraints(s, s->z[0]);
            s->iterOrder = 0;
            s->odeHStatus = H_CONST;
        }
        s->h = s->hNextStep;
        s->hNordsieck = s->hNextStep;
        if (fabs(s->h) < EPS_F * s->min_step_size) { s->status = ST_SMALLH; return; }
        s->t = s->tInMeshPoint + s->h;
        if (s->t == s->tInMeshPoint) { s->status = ST_SMALLH; return; }
        multi_value_step(s);
        if (s->convergenceStatus == CONV_FAIL) {
            s->iterConvergenceFailure++;
            s->numberOfConvergenceFailure++;
            if (s->iterConvergenceFailure >= 50) { s->status = ST_CONVFAIL; return; }
            if (fabs(s->h - s->min_step_size) <= EPS_D * fabs(s->min_step_size)) { s->status = ST_SMALLH; return; }
            convergence_failure(s);
        } else {
            s->numberOfConvergenceSuccess++;
            double hsc = s->Ep[s->p] * wrms(ne, s->v[s->p + 1], s->ooe);
            if (hsc > 1.) {
                s->iterErrorFailure++;
                s->numberOfErrorCheckFailure++;
                if (s->iterErrorFailure >= 50) { s->status = ST_ERRFAIL; return; }
                if (fabs(s->h - s->min_step_size) <= EPS_D * fabs(s->min_step_size)) { s->status = ST_SMALLH; return; }
                s->hScaleMax = 1.;
                hsc = 0.8 * pow(1. / hsc, 1. / (double)(s->p + 1));
                s->hScale = hsc;
                if (s->p > MIN_ORDER) {
                    double hm = s->Ep[s->p - 1] * wrms(ne, s->v[s->p], s->ooe);
                    hm = 1. / (hm + EPS_F);
                    hm = 0.8 * pow(hm, 1. / (double)s->p);
                    if (hm > hsc) { s->hScale = hm; s->p -= 1; s->odeOrderStatus = O_DEC; }
                }
                s->hScale = fmin(s->hScale, 0.5);
                s->hScale = fmax(s->hScale, s->min_step_size / fabs(s->h));
                s->hScale = fmax(s->hScale, 1.e-3);
                if (s->p == 1 && s->hScale < 0.1 && s->numberOfSteps > 10 && s->tInMeshPoint != s->tStabilize) {
                    eval(s, s->z[0], s->va);
                    s->numberOfFunctionCalls++;
                    for (int q = 0; q < ne; q++) s->z[1][q] = s->va[q] * s->h;
                    s->tStabilize = s->tInMeshPoint;
                }
                s->odeHStatus = H_DEC;
                s->hNextStep *= s->hScale;
                s->numberOfDecreasedSteps++;
            } else {
                s->numberOfErrorCheckSuccess++;
                s->hNordsieck = s->h; s->hPreviousStep = s->h; s->hNextStep = s->h;
                s->orderInNextStep = s->p;
                s->odeOrderStatus = O_CONST; s->odeHStatus = H_CONST;
                convergence_rate(s);
                if (s->iterErrorFailure > 0 || s->iterConvergenceFailure > 0) {
                    s->iterErrorFailure = 0; s->iterConvergenceFailure = 0;
                } else {
                    s->iterOrder++;
                    if (s->iterOrder > (unsigned)s->p + 1) new_order_new_h(s);
                }
                for (int i = 0; i <= s->p + 2; i++) memcpy(s->z[i], s->v[i], sizeof(double) * ne);
                check_constraints(s, s->z[0]);
                if (s->odeOrderStatus == O_CONST && s->odeHStatus == H_CONST) s->numberOfStepsWithoutChanges++;
                break;
            }
        }
    }
}

/* Interpolation MultiValueSolver.hpp:688-706 */
static void interpolation(bdf_t *s) {
    const double hp = (s->tOut - s->t) / s->hNordsieck;
    double hr = 1.;
    memcpy(s->y, s->z[0], sizeof(double) * s->ne);
    if (hp == 0.) return;
    for (int i = 1; i <= s->p; i++) {
        hr *= hp;
        for (int q = 0; q < s->ne; q++) { s->va[q] = s->z[i][q] * hr; s->y[q] += s->va[q]; }
    }
    check_constraints(s, s->y);
}

/* MultiValueSolver::Solve MultiValueSolver.hpp:424-617 (first call; tolerances
 * scalar; no strict tmax; no y' stop criterion) */
static int bdf_solve(bdf_t *s, double tOut) {
    const int ne = s->ne;
    s->tOut = tOut;
    if (illegal_constraints(s, s->y0) != 0) { s->status = ST_ILLEGAL; return s->status; }
    s->first_step_size = initial_step_size(s);
    memcpy(s->z[0], s->y0, sizeof(double) * ne);
    for (int i = 0; i < ne; i++) s->z[1][i] = s->f[i] * s->first_step_size;
    s->hPreviousStep = s->hNordsieck = s->hNextStep = s->first_step_size;
    s->status = ST_CONT;
    for (unsigned k = 1; k <= s->max_steps; k++) {
        one_over_epsilon(s, s->z[0]);
        const double tolSafe = EPS_D * wrms(ne, s->z[0], s->ooe);
        if (tolSafe > 1.) { s->status = ST_STRICT; return s->status; }
        s->tInMeshPoint = s->t;
        driver_step(s);
        if (s->status <= 0) { s->t = s->tInMeshPoint; return s->status; }
        s->numberOfSteps++;
        if (s->maxOrderUsed < (unsigned)s->p) s->maxOrderUsed = s->p;
        if (s->hMaxUsed < s->h) s->hMaxUsed = s->h;
        if (s->hMinUsed > s->h) s->hMinUsed = s->h;
        if (s->tOut == s->t) { memcpy(s->y, s->z[0], sizeof(double) * ne); return s->status; }
        if (s->tOut < s->t) {
            interpolation(s);
            eval(s, s->y, s->va);          /* printResults_ defaults to true (:101) */
            return s->status;
        }
    }
    s->status = ST_MAXSTEPS;
    return s->status;
}

static void fill_counters(const bdf_t *s, long rhs_total, int32_t *c) {
    if (!c) return;
    c[0] = (int32_t)s->numberOfSteps;
    c[1] = (int32_t)s->numberOfFunctionCalls;
    c[2] = (int32_t)s->numberOfFunctionCallsForJacobian;
    c[3] = (int32_t)s->numberOfJacobians;
    c[4] = (int32_t)s->numberOfMatrixFactorizations;
    c[5] = (int32_t)s->numberOfLinearSystemSolutions;
    c[6] = (int32_t)s->numberOfConvergenceFailure;
    c[7] = (int32_t)s->numberOfErrorCheckFailure;
    c[8] = (int32_t)s->maxOrderUsed;
    c[9] = (int32_t)rhs_total;
    c[10] = (int32_t)s->clamps;
    c[11] = (int32_t)s->newton_iters;
}

typedef struct { lsm_oracle_rhs_fn f; void *ctx; long n; } counting_t;
static void counting_rhs(void *v, const double *y, double *f) { counting_t *c = v; c->n++; c->f(c->ctx, y, f); }

int lsm_oracle_bdf(int n, lsm_oracle_rhs_fn f, void *ctx, const lsm_oracle_bdf_options *o, double t0,
                   const double *y0, double tOut, double *yout, int32_t *counters) {
    bdf_t s;
    counting_t cc = {f, ctx, 0};
    bdf_alloc(&s, n);
    s.rhs = counting_rhs; s.ctx = &cc;
    s.rel_tol = o->rel_tol; s.abs_tol = o->abs_tol; s.full_pivoting = o->full_pivoting;
    if (o->max_steps > 0) s.max_steps = (unsigned)o->max_steps;
    s.min_c = o->use_min; s.max_c = o->use_max; s.ymin = o->ymin; s.ymax = o->ymax;
    set_initial_conditions(&s, t0, y0);
    const int st = bdf_solve(&s, tOut);
    memcpy(yout, st > 0 ? s.y : y0, sizeof(double) * n);
    fill_counters(&s, cc.n, counters);
    bdf_free(&s);
    return st;
}

/* ------------------------------------------------------------------------- */
/* Per-cell driver: chemistry_DI.H:76-331 (OpenSMOKE branch, const-P)        */
/* ------------------------------------------------------------------------- */
typedef struct {
    mech_t M;
    reactor_t R;
    bdf_t s;
    double *ymin, *ymax, *y0;
} cell_ctx_t;

static void cell_ctx_init(cell_ctx_t *c, const lsm_mechanism *m, const lsm_oracle_options *o) {
    const int ns = m->ns, ne = ns + 1;
    mech_init(&c->M, m);
    c->R.M = &c->M; c->R.P = o->pressure; c->R.energy = o->energy_equation; c->R.QR = 0.; c->R.nrhs = 0;
    c->R.cv = o->constant_volume ? 1 : 0; c->R.rho0 = 0.;
    kin_ws_alloc(&c->R.w, ns, m->nr);
    c->ymin = malloc(8 * ne); c->ymax = malloc(8 * ne); c->y0 = malloc(8 * ne);
    bdf_alloc(&c->s, ne);
    for (int i = 0; i < ne; i++) { c->ymin[i] = 0.; c->ymax[i] = 1.; }   /* chemistry_DI.H:82-83 */
    c->ymin[ns] = 200.; c->ymax[ns] = 6000.;
}

static void cell_ctx_free(cell_ctx_t *c) {
    kin_ws_free(&c->R.w); mech_free(&c->M); free(c->ymin); free(c->ymax); free(c->y0);
    bdf_free(&c->s);
}

/* normalizeMassFractions, LS/utilities.H:153-203 (no abort: callers get status) */
static void normalize_mass_fractions(double *y, int ns) {
    double sum = 0.;
    for (int i = 0; i < ns; i++) sum += y[i];
    for (int i = 0; i < ns; i++) y[i] /= sum;
}

/* Strang "compact" sources of the final state (chemistry_DI.H:236-282):
 * RR_i = rho (yf_i - Y_i)/dt, RT = rho cp (Tf - T)/dt with rho and cp of the final
 * state at the thermodynamic pressure, Y and T the cell's input fields; zero for
 * dt <= 1e-14 (and RT zero without the energy equation). */
static void compact_sources(cell_ctx_t *c, const lsm_oracle_options *o, double dt, const double *yin, double *yout) {
    const mech_t *M = &c->M;
    const int ns = M->ns;
    const double Tf = yout[ns];
    double u = 0.;
    for (int i = 0; i < ns; i++) u += yout[i] / M->MW[i];
    const double MW = 1. / u;
    thermo_eval(M, Tf, c->R.w.cp, c->R.w.h, c->R.w.s);
    double cpd = 0.;
    for (int i = 0; i < ns; i++) cpd += c->R.w.cp[i] * ((MW * yout[i]) / M->MW[i]);
    const double cp = cpd * LSM_R_J_KMOL / MW;
    const double rho = o->pressure * MW / LSM_R_J_KMOL / Tf;
    const int ok = dt > 1.e-14;
    for (int i = 0; i < ns; i++) yout[i] = ok ? rho * (yout[i] - yin[i]) / dt : 0.;
    yout[ns] = (ok && o->energy_equation) ? rho * cp * (Tf - yin[ns]) / dt : 0.;
}

static int integrate_cell(cell_ctx_t *c, const lsm_oracle_options *o, double dt, double rho_cell,
                          const double *yin, double *yout, double *q, int32_t *counters) {
    const int ns = c->M.ns, ne = ns + 1;
    int status = 0;
    if (yin[ns] > o->min_temperature) {                         /* chemistry_DI.H:97 */
        double sum = 0.;                                        /* :101-128 */
        for (int i = 0; i < ns; i++) { c->y0[i] = yin[i] < 0. ? 0. : yin[i]; sum += c->y0[i]; }
        for (int i = 0; i < ns; i++) c->y0[i] /= sum;
        c->y0[ns] = yin[ns];
        if (c->R.cv) {                                          /* SetReactor(V, P, rho), :173 */
            double u = 0.;
            for (int i = 0; i < ns; i++) u += c->y0[i] / c->M.MW[i];
            c->R.rho0 = rho_cell > 0. ? rho_cell : o->pressure * (1. / u) / LSM_R_J_KMOL / c->y0[ns];
        }
        bdf_fresh(&c->s);                                       /* fresh solver per cell */
        c->s.rhs = reactor_rhs; c->s.ctx = &c->R;
        c->s.rel_tol = o->rel_tol; c->s.abs_tol = o->abs_tol; c->s.full_pivoting = o->full_pivoting;
        if (o->max_steps > 0) c->s.max_steps = (unsigned)o->max_steps;
        if (o->jac_increment_scale > 0.) c->s.jac_scale = o->jac_increment_scale;
        c->M.rev_sum = o->reverse_sums ? 1 : 0;
        c->s.min_c = c->s.max_c = 1; c->s.ymin = c->ymin; c->s.ymax = c->ymax;
        c->R.nrhs = 0;
        set_initial_conditions(&c->s, o->t0, c->y0);            /* :138 */
        status = bdf_solve(&c->s, o->t0 + dt);                  /* :157 */
        if (status > 0) {
            memcpy(yout, c->s.y, 8 * ne);
            if (q) *q = c->R.QR;                                /* :168 */
        } else {
            memcpy(yout, c->y0, 8 * ne);                        /* failure: keep the input state */
        }
        fill_counters(&c->s, c->R.nrhs, counters);
    } else {
        memcpy(yout, yin, 8 * ne);
        if (counters) memset(counters, 0, sizeof(int32_t) * LSM_ORACLE_NCOUNTERS);
    }
    normalize_mass_fractions(yout, ns);                         /* :219 */
    if (!o->energy_equation) yout[ns] = yin[ns];                /* :231-233 */
    if (o->strang_compact) compact_sources(c, o, dt, yin, yout);
    return status;
}

int lsm_oracle_integrate_cell(const lsm_mechanism *m, const lsm_oracle_options *o, double dt,
                              const double *y_in, double *y_out, double *q, int32_t *counters) {
    cell_ctx_t c;
    cell_ctx_init(&c, m, o);
    const int st = integrate_cell(&c, o, dt, o->rho ? o->rho[0] : 0., y_in, y_out, q, counters);
    cell_ctx_free(&c);
    return st;
}

typedef struct {
    const lsm_mechanism *m;
    const lsm_oracle_options *o;
    int64_t n;
    const double *dt, *yin;
    double *yout, *q;
    int32_t *status, *counters;
    int64_t next;
    pthread_mutex_t lock;
} pool_t;

static void *worker(void *arg) {
    pool_t *P = arg;
    const int ns = P->m->ns, ne = ns + 1;
    cell_ctx_t c;
    cell_ctx_init(&c, P->m, P->o);
    double *yi = malloc(8 * ne), *yo = malloc(8 * ne);
    for (;;) {
        pthread_mutex_lock(&P->lock);
        const int64_t b = P->next;
        P->next += 64;
        pthread_mutex_unlock(&P->lock);
        if (b >= P->n) break;
        const int64_t e = b + 64 < P->n ? b + 64 : P->n;
        for (int64_t cell = b; cell < e; cell++) {
            for (int i = 0; i < ne; i++) yi[i] = P->yin[(int64_t)i * P->n + cell];
            double qv = P->q ? P->q[cell] : 0.;
            const int st = integrate_cell(&c, P->o, P->dt[cell], P->o->rho ? P->o->rho[cell] : 0., yi, yo, &qv,
                                          P->counters ? P->counters + cell * LSM_ORACLE_NCOUNTERS : NULL);
            for (int i = 0; i < ne; i++) P->yout[(int64_t)i * P->n + cell] = yo[i];
            if (P->q) P->q[cell] = qv;
            if (P->status) P->status[cell] = st;
        }
    }
    free(yi); free(yo);
    cell_ctx_free(&c);
    return NULL;
}

void lsm_oracle_chemistry_step(const lsm_mechanism *m, const lsm_oracle_options *o, int64_t n,
                               const double *dt, const double *y_in, double *y_out, int32_t *status,
                               double *q, int32_t *counters, int n_threads) {
    pool_t P;
    memset(&P, 0, sizeof(P));
    P.m = m; P.o = o; P.n = n; P.dt = dt; P.yin = y_in; P.yout = y_out;
    P.q = q; P.status = status; P.counters = counters;
    pthread_mutex_init(&P.lock, NULL);
    if (n_threads < 1) n_threads = 1;
    if (n_threads == 1) { worker(&P); }
    else {
        pthread_t *th = malloc(sizeof(pthread_t) * n_threads);
        for (int i = 0; i < n_threads; i++) pthread_create(&th[i], NULL, worker, &P);
        for (int i = 0; i < n_threads; i++) pthread_join(th[i], NULL);
        free(th);
    }
    pthread_mutex_destroy(&P.lock);
}

/* ------------------------------------------------------------------------- */
/* Transport: TransportPropertiesMap_CHEMKIN.hpp                              */
/* ------------------------------------------------------------------------- */
void lsm_oracle_transport(const lsm_transport *tr, double T, double P, const double *x, double *mu,
                          double *lambda, double *dmix, double *teta) {
    const int ns = tr->ns;
    const double logT = log(T), logT2 = logT * logT, logT3 = logT * logT2;
    double *lam = malloc(sizeof(double) * ns), *eta = malloc(sizeof(double) * ns);
    double *sumK = malloc(sizeof(double) * ns), *xc = malloc(sizeof(double) * ns);
    double *sdc = malloc(sizeof(double) * ns);
    /* lambda(), eta(): :690-766 */
    for (int j = 0; j < ns; j++) {
        const double *k = tr->lambda_fit + 4 * j, *e = tr->eta_fit + 4 * j;
        double s1 = k[0];
        s1 += k[1] * logT; s1 += k[2] * logT2; s1 += k[3] * logT3;
        lam[j] = exp(s1);
        double s2 = e[0];
        s2 += e[1] * logT; s2 += e[2] * logT2; s2 += e[3] * logT3;
        eta[j] = exp(s2);
    }
    /* lambdaMix, :947-957 (Mathur, Todor, Saxena) */
    if (lambda) {
        double sum1 = 0., sum2 = 0.;
        for (int i = 0; i < ns; i++) { sum1 += x[i] * lam[i]; sum2 += x[i] / lam[i]; }
        *lambda = 0.50 * (sum1 + 1. / sum2);
    }
    /* etaMix, :959-1030 */
    if (mu) {
        const double *M = tr->mw;
        if (tr->viscosity_model == LSM_VISCOSITY_WILKE) {
            for (int k = 0; k < ns; k++) sumK[k] = x[k];
            double *sq = malloc(sizeof(double) * ns), *usq = malloc(sizeof(double) * ns);
            for (int k = 0; k < ns; k++) { sq[k] = sqrt(eta[k]); usq[k] = 1. / sq[k]; }
            for (int k = 0; k < ns; k++)
                for (int j = k + 1; j < ns; j++) {
                    /* CompleteInitialization, :655-670 */
                    const double sup = 1. / sqrt(8. * (1. + M[k] / M[j]));
                    const double inf = 1. / sqrt(8. * (1. + M[j] / M[k]));
                    const double r14 = sqrt(sqrt(M[j] / M[k]));
                    const double d = sq[k] * usq[j] * r14;
                    sumK[k] += x[j] * sup * (1. + d) * (1. + d);
                    sumK[j] += x[k] * inf * (1. + 1. / d) * (1. + 1. / d);
                }
            double m = 0.;
            for (int k = 0; k < ns; k++) m += x[k] * eta[k] / sumK[k];
            *mu = m;
            free(sq); free(usq);
        } else if (tr->viscosity_model == LSM_VISCOSITY_HERNING) {
            double sum = 0., m = 0.;
            for (int k = 0; k < ns; k++) sum += x[k] * sqrt(M[k]);
            for (int k = 0; k < ns; k++) m += x[k] * eta[k] * sqrt(M[k]);
            *mu = m / sum;
        } else {
            double sum1 = 0., sum2 = 0.;
            for (int i = 0; i < ns; i++) { sum1 += x[i] * eta[i]; sum2 += x[i] / eta[i]; }
            *mu = 0.50 * (sum1 + 1. / sum2);
        }
    }
    /* gamma() and gammaMix, :798-859, 1032-1059 */
    if (dmix) {
        const double threshold = 1.e-14, sum_threshold = 1. + threshold * (double)ns;
        const double P_bar = P / 100000.;
        for (int k = 0; k < ns; k++) sdc[k] = 0.;
        for (int i = 0; i < ns; i++) xc[i] = (x[i] + threshold) / sum_threshold;
        double MWmix = 0.;
        for (int i = 0; i < ns; i++) MWmix += xc[i] * tr->mw[i];
        const double *d = tr->gamma_fit;
        for (int k = 0; k < ns; k++)
            for (int j = k + 1; j < ns; j++) {
                double sum = d[0];
                sum += d[1] * logT; sum += d[2] * logT2; sum += d[3] * logT3;
                const double D = P_bar / exp(sum);
                sdc[j] += xc[k] * D;
                sdc[k] += xc[j] * D;
                d += 4;
            }
        for (int k = 0; k < ns; k++) dmix[k] = (MWmix - xc[k] * tr->mw[k]) / (MWmix * sdc[k]);
    }
    /* teta() and tetaMix, :768-796, 1103-1117 */
    if (teta) {
        const double T2 = T * T, T3 = T * T2;
        for (int j = 0; j < ns; j++) teta[j] = 0.;
        for (int q = 0; q < tr->nteta; q++) {
            const int i = tr->teta_sp[q];
            const double *f = tr->teta_fit + 4 * (size_t)q * ns;
            for (int j = 0; j < ns; j++) {
                double sum = f[4 * j];
                sum += f[4 * j + 1] * T; sum += f[4 * j + 2] * T2; sum += f[4 * j + 3] * T3;
                teta[i] += sum * x[i] * x[j];
            }
        }
    }
    free(lam); free(eta); free(sumK); free(xc); free(sdc);
}

/* properties.H:79-192, one cell */
void lsm_oracle_refresh_cell(const lsm_mechanism *m, const lsm_transport *tr, const lsm_transport_options *o,
                             double P, const double *y, double *mu, double *lambda, double *dmix, double *teta) {
    const int ns = m->ns;
    const double T = y[ns];
    double *x = malloc(sizeof(double) * ns), *cpR = malloc(sizeof(double) * ns);
    double *hRT = malloc(sizeof(double) * ns), *sR = malloc(sizeof(double) * ns);
    double *D = malloc(sizeof(double) * ns), *th = malloc(sizeof(double) * ns);
    /* MoleFractions_From_MassFractions (ThermodynamicsMap_CHEMKIN.hpp:429-434) */
    double u = 0.;
    for (int k = 0; k < ns; k++) u += y[k] / m->nasa[18 * k + 17];
    const double MW = 1. / u;
    for (int k = 0; k < ns; k++) x[k] = MW * y[k] / m->nasa[18 * k + 17];
    const double cTot = P / LSM_R_J_KMOL / T, psi = cTot * MW / P;          /* :97-98 */
    const int lewis = o && o->diffusivity_model == LSM_DIFFUSIVITY_LEWIS_NUMBERS;
    const int soret = o && o->soret;
    double mu_c, lam_c;
    lsm_oracle_transport(tr, T, P, x, &mu_c, &lam_c, lewis ? NULL : D, th);
    /* cpMolar_Mixture_From_MoleFractions / MW (:109-113) */
    lsm_oracle_species_thermo(m, T, cpR, hRT, sR);
    double cp = 0.;
    for (int k = 0; k < ns; k++) cp += cpR[k] * x[k];
    cp = cp * LSM_R_J_KMOL / MW;
    if (o && o->simplified) {                                              /* :130-139 */
        mu_c = 1.8405e-5 * pow(T / 300., 0.6759);
        lam_c = mu_c * cp / 0.7;
    }
    if (lewis) {                                                           /* :153-164 */
        const double coefficient = lam_c / (P * psi) / cp;
        for (int k = 0; k < ns; k++) D[k] = coefficient / o->lewis[k];
    }
    if (o && o->soot_reference >= 0) {                                     /* :166-177 */
        const double Dref = D[o->soot_reference];
        for (int i = 0; i < o->n_soot; i++) D[o->soot_sp[i]] = Dref * o->soot_corr[i];
    }
    if (soret)                                                             /* :180-192 */
        for (int q = 0; q < tr->nteta; q++) {
            const int k = tr->teta_sp[q];
            th[k] = D[k] * th[k] * m->nasa[18 * k + 17] / MW;
        }
    if (mu) *mu = mu_c;
    if (lambda) *lambda = lam_c;
    if (dmix) memcpy(dmix, D, sizeof(double) * ns);
    if (teta) memcpy(teta, th, sizeof(double) * ns);
    free(x); free(cpR); free(hRT); free(sR); free(D); free(th);
}
