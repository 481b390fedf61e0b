// lsm_kernels.hip - CDNA4 (gfx950) kernels of the chemistry step.
//
//  * team kernel  : one cell per 64-lane wavefront, persistent work queue.  Lane
//                   k owns species k (element k of every vector), the reaction
//                   loop is strided over lanes, the Newton matrix is factorised
//                   in registers (lane i holds row i) with the pivot row
//                   broadcast by v_readlane, the Jacobian lives in LDS.
//  * plain kernel : one cell per thread, all per-cell state in HBM with the
//                   cell index on the fast axis (coalesced across the wave).
// Both run the same Gear/BDF driver (lsm_bdf.hpp) and the same reaction-rate
// body, which restates KineticsMap_CHEMKIN::KineticConstants/ReactionRates
// (maps/KineticsMap_CHEMKIN.hpp:1192-1386, 1683-1768) for one reaction.
#include <hip/hip_runtime.h>

#include <math.h>

#include <utility>

#include "../../include/lsm_mechanism.h"
#include "../../include/lsm_transport.h"
#include "lsm_internal.hpp"

// The file is compiled as several translation units in parallel (laminarsmoke_amd/build.py):
// LSM_PART 1-5 each instantiate the team kernel for one register-array size (NMAX 16, 32,
// 48, 56, 64), LSM_PART 6 holds every other kernel and the host launchers; LSM_PART 0 (the
// default) is all of it in one unit.
#ifndef LSM_PART
#define LSM_PART 0
#endif
#define LSM_IN_PART(p) (LSM_PART == 0 || LSM_PART == (p))

namespace lsm {

constexpr int LSM_NCOUNTERS_DEV = 12;
constexpr double LN10 = 2.302585092994045684;
constexpr double RGAS = LSM_R_J_KMOL;
constexpr double NEG_INF = -__builtin_huge_val();
constexpr double POS_INF = __builtin_huge_val();

// Optional phase timers (-DLSM_PHASE_TIMERS): cycles spent in the RHS, the
// Jacobian (including its RHS calls), the factorisation, the solves and J*v,
// summed over waves into KParams::phase_cycles[0..4]; [5] whole cells, [6] cells;
// team RHS parts: [7] species and sums, [8] reactions, [9] assembly; wide solves: [7] FP32
// triangular solves, [8] FP64 residual mat-vecs, [9] refinement sweeps (a count).
#ifdef LSM_PHASE_TIMERS
#define LSM_TIC(v) const unsigned long long v = clock64()
#define LSM_TOC(i, v) (tp[i] += clock64() - (v))
#else
#define LSM_TIC(v)
#define LSM_TOC(i, v)
#endif

// ---------------------------------------------------------------------------
// wavefront helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double uniform_d(double x) {
    const long long v = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readfirstlane((int)(v & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double readlane_d(double x, int l) {
    const long long v = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// 64-bit DPP move of both halves; lanes without a source (or outside row_mask) get 0.
template <int CTRL, int ROWM>
__device__ __forceinline__ long long dpp0_i64(long long v) {
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(v & 0xffffffffll), CTRL, ROWM, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(v >> 32), CTRL, ROWM, 0xf, false);
    return ((long long)hi << 32) | (unsigned)lo;
}
template <int CTRL, int ROWM>
__device__ __forceinline__ double dpp0_d(double v) {
    return __longlong_as_double(dpp0_i64<CTRL, ROWM>(__double_as_longlong(v)));
}

// Inclusive wavefront scan with row_shr:1,2,4,8 then row_bcast:15,31 (gfx9 DPP);
// lane 63 ends with the total, which is broadcast.  Fixed summation order.
__device__ __forceinline__ double wave_sum(double x) {
    x += dpp0_d<0x111, 0xf>(x);
    x += dpp0_d<0x112, 0xf>(x);
    x += dpp0_d<0x114, 0xf>(x);
    x += dpp0_d<0x118, 0xf>(x);
    x += dpp0_d<0x142, 0xa>(x);
    x += dpp0_d<0x143, 0xc>(x);
    return readlane_d(x, 63);
}

// Two independent sums in one pass (interleaved for latency).
__device__ __forceinline__ void wave_sum2(double &a, double &b) {
    a += dpp0_d<0x111, 0xf>(a); b += dpp0_d<0x111, 0xf>(b);
    a += dpp0_d<0x112, 0xf>(a); b += dpp0_d<0x112, 0xf>(b);
    a += dpp0_d<0x114, 0xf>(a); b += dpp0_d<0x114, 0xf>(b);
    a += dpp0_d<0x118, 0xf>(a); b += dpp0_d<0x118, 0xf>(b);
    a += dpp0_d<0x142, 0xa>(a); b += dpp0_d<0x142, 0xa>(b);
    a += dpp0_d<0x143, 0xc>(a); b += dpp0_d<0x143, 0xc>(b);
    a = readlane_d(a, 63);
    b = readlane_d(b, 63);
}

__device__ __forceinline__ unsigned long long umax64(unsigned long long a, unsigned long long b) {
    return a > b ? a : b;
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long x) {
    x = umax64(x, (unsigned long long)dpp0_i64<0x111, 0xf>((long long)x));
    x = umax64(x, (unsigned long long)dpp0_i64<0x112, 0xf>((long long)x));
    x = umax64(x, (unsigned long long)dpp0_i64<0x114, 0xf>((long long)x));
    x = umax64(x, (unsigned long long)dpp0_i64<0x118, 0xf>((long long)x));
    x = umax64(x, (unsigned long long)dpp0_i64<0x142, 0xa>((long long)x));
    x = umax64(x, (unsigned long long)dpp0_i64<0x143, 0xc>((long long)x));
    const long long v = (long long)x;
    const int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffffll), 63);
    const int hi = __builtin_amdgcn_readlane((int)(v >> 32), 63);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}

// Wavefront maximum of non-negative doubles (lanes without a DPP source read 0), broadcast.
__device__ __forceinline__ double wave_max_nonneg(double x) {
    x = fmax(x, dpp0_d<0x111, 0xf>(x));
    x = fmax(x, dpp0_d<0x112, 0xf>(x));
    x = fmax(x, dpp0_d<0x114, 0xf>(x));
    x = fmax(x, dpp0_d<0x118, 0xf>(x));
    x = fmax(x, dpp0_d<0x142, 0xa>(x));
    x = fmax(x, dpp0_d<0x143, 0xc>(x));
    return readlane_d(x, 63);
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// c^nu by stoichiometric class (StoichiometricMap.hpp:537-552)
__device__ __forceinline__ double conc_pow(double c, double nu) {
    if (nu == 1.) return c;
    if (nu == 2.) return c * c;
    if (nu == 3.) return (c * c) * c;
    if (nu == 0.5) return sqrt(c);
    return pow(c, nu);
}

// A factor c^lambda of a non-elementary side (an order below 1 on that side):
// StoichiometricMap::ProductOfConcentrationsForNonElementaryReactions
// (StoichiometricMap.hpp:617-654) - c^lambda for lambda >= 1; below 1 a tanh blend of
// (c + m/1e9)^lambda and the line c*^(lambda-1) c through 0, c* = 1e-8 kmol/m3.
__device__ __forceinline__ double order_pow(double c, double lambda) {
    if (lambda >= 1.) return pow(c, lambda);
    const double Cs = 1.e-8, ALFA = 1.e-5;
    const double H = 1.50 * log(ALFA / (1. - ALFA));
    const double K = 2.00 * log((1. - ALFA) / ALFA) / Cs;
    const double m = (tanh(K * c + H) + 1.) / 2.;
    return m * pow(c + m / 1.e9, lambda) + (1. - m) * pow(Cs, lambda - 1.) * c;
}

// d order_pow / dc
__device__ __forceinline__ double order_pow_der(double c, double lambda) {
    if (lambda >= 1.) return lambda == 1. ? 1. : ((c > 0. || lambda > 1.) ? lambda * pow(c, lambda - 1.) : 0.);
    const double Cs = 1.e-8, ALFA = 1.e-5;
    const double H = 1.50 * log(ALFA / (1. - ALFA));
    const double K = 2.00 * log((1. - ALFA) / ALFA) / Cs;
    const double th = tanh(K * c + H);
    const double m = (th + 1.) / 2., dm = 0.5 * K * (1. - th * th);
    const double a = c + m / 1.e9;
    const double csl = pow(Cs, lambda - 1.);
    return dm * pow(a, lambda) + m * lambda * pow(a, lambda - 1.) * (1. + dm / 1.e9) - dm * csl * c + (1. - m) * csl;
}

// Forward constant of a PLOG reaction at pressure P [Pa]
// (PressureLogarithmicRateExpression::KineticConstant,
// kernel/kinetics/PressureLogarithmicRateExpression.hpp:104-119): n levels
// lv[4i..4i+3] = ln p_i, lnA_i, beta_i, E_i/R with p_i increasing; the end level
// outside [p_0, p_{n-1}], else ln k linear in ln P between the two levels around P.
// Returns ln k; dlnk = d ln k / dT at fixed P.
__device__ __forceinline__ double plog_lnk(const double *lv, int n, double P, double logT, double uT,
                                           double &dlnk) {
    const double lnP = log(P);
    int i = 0;
    double w = 0.;
    if (lnP <= lv[0]) {
        i = 0;
    } else if (lnP >= lv[4 * (n - 1)]) {
        i = n - 1;
    } else {
        while (i < n - 2 && !(lnP < lv[4 * (i + 1)])) i++;
        w = (lnP - lv[4 * i]) / (lv[4 * (i + 1)] - lv[4 * i]);
    }
    const double *a = lv + 4 * i;
    const double lnkA = a[1] + a[2] * logT - a[3] * uT;
    const double dA = (a[2] + a[3] * uT) * uT;
    if (w == 0.) { dlnk = dA; return lnkA; }
    const double *b = a + 4;
    const double lnkB = b[1] + b[2] * logT - b[3] * uT;
    dlnk = dA + ((b[2] + b[3] * uT) * uT - dA) * w;
    return lnkA + (lnkB - lnkA) * w;
}

// Forward constant of a Chebyshev reaction at (T, P)
// (ChebyshevPolynomialRateExpression::KineticConstant,
// kernel/kinetics/ChebyshevPolynomialRateExpression.hpp:95-116): d = [N, M, Tmin, Tmax,
// log10 Pmin, log10 Pmax, conv, a_00 .. a_(N-1)(M-1)], k = conv 10^(sum a_nm phi_n(T~)
// phi_m(P~)).  phi_n(x) = cos(n acos x) is evaluated by the recurrence
// phi_(n+1) = 2 x phi_n - phi_(n-1) (the same polynomial, T_n(cos t) = cos(n t)); outside
// [-1, 1] the reference's acos is undefined, and k is NaN here as well.
// dlnk = d ln k / dT at fixed P (T_n' from its own recurrence).
__device__ __forceinline__ double cheb_k(const double *d, double uT, double P, double &dlnk) {
    const int N = (int)d[0], M = (int)d[1];
    const double Tt = (2. * uT - 1. / d[2] - 1. / d[3]) / (1. / d[3] - 1. / d[2]);
    const double Pt = (2. * log10(P) - d[4] - d[5]) / (d[5] - d[4]);
    if (!(fabs(Tt) <= 1.) || !(fabs(Pt) <= 1.)) { dlnk = NAN; return NAN; }
    const double *a = d + 7;
    double sum = 0., dsum = 0.;
    double t0 = 1., t1 = Tt, dt0 = 0., dt1 = 1.;
    for (int n = 0; n < N; n++) {
        const double tn = n == 0 ? t0 : t1, dtn = n == 0 ? dt0 : dt1;
        double p0 = 1., p1 = Pt;
        for (int m = 0; m < M; m++) {
            const double pm = m == 0 ? p0 : p1;
            sum += a[n * M + m] * tn * pm;
            dsum += a[n * M + m] * dtn * pm;
            if (m > 0) { const double p2 = 2. * Pt * p1 - p0; p0 = p1; p1 = p2; }
        }
        if (n > 0) {
            const double t2 = 2. * Tt * t1 - t0, dt2 = 2. * t1 + 2. * Tt * dt1 - dt0;
            t0 = t1; t1 = t2; dt0 = dt1; dt1 = dt2;
        }
    }
    dlnk = LN10 * dsum * (-2. * uT * uT / (1. / d[3] - 1. / d[2]));
    return exp(LN10 * sum) * d[6];
}

// Net rate of reaction j [kmol/m3/s].  C, H, S return c_k, h_k/RT, s_k/R; P is the
// pressure the kinetics see (it only enters PLOG constants).  XK: the mechanism has
// explicitly reversible, chemically activated or PLOG reactions (without them the
// code for those forms is not compiled in).
template <bool XK, class FC, class FH, class FS>
__device__ __forceinline__ double reaction_rate(const DevMech &M, int j, double T, double logT, double uT,
                                                double logPRT, double P, double cTot, FC C, FH H, FS S) {
    // Arrhenius (hpp:1198-1212); PLOG constants replace it (hpp:1342-1349)
    const int info = M.rinfo[j];
    double kf;
    if (XK && M.rpl[j] >= 0) {
        const int s = M.rpl[j];
        double d;
        kf = exp(plog_lnk(M.pl_lev + 4 * M.pl_ptr[s], M.pl_ptr[s + 1] - M.pl_ptr[s], P, logT, uT, d));
    } else if (XK && M.rch[j] >= 0) {    // Chebyshev (hpp:1333-1340)
        double d;
        kf = cheb_k(M.ch_dat + M.ch_ptr[M.rch[j]], uT, P, d);
    } else {
        kf = exp(M.rA[j] + M.rB[j] * logT - M.rE[j] * uT);
        if (info & LSM_RXN_NEGATIVE_A) kf = -kf;
    }
    double kmod = kf;
    // third body (hpp:1388-1398, 1703-1708)
    const int tb = M.rtb[j];
    if (tb >= 0) {
        double Me = cTot;
        for (int q = M.tb_ptr[tb]; q < M.tb_ptr[tb + 1]; q++) Me += C(M.tb_sp[q]) * M.tb_eff[q];
        kmod *= Me;
    }
    // fall-off (hpp:1256-1290, 1496-1550, 1710-1716)
    const int fo = M.rfo[j];
    if (fo >= 0) {
        const double *fp = M.fo_par + 8 * fo;
        const double kinf = exp(fp[0] + fp[1] * logT - fp[2] * uT);
        double Mc;
        if (M.fo_single[fo] < 0) {
            Mc = cTot;
            for (int q = M.fo_ptr[fo]; q < M.fo_ptr[fo + 1]; q++) Mc += C(M.fo_sp[q]) * M.fo_eff[q];
        } else {
            Mc = C(M.fo_single[fo]) + 1.e-16;
        }
        const double Pr = kf * Mc / kinf;
        const int form = M.fo_type[fo] & 3;
        const bool cabr = XK && (M.fo_type[fo] & LSM_FALLOFF_CHEMACT);
        double wF = 1.;
        if (form == LSM_FALLOFF_TROE) {
            const double a = fp[3];
            double F = (1. - a) * exp(-T / fp[4]) + a * exp(-T / fp[5]);
            if (fp[6] != 0.) F += exp(-fp[6] / T);
            const double lf = (F < 1.e-300) ? -300. : log10(F);
            double q;
            if (Pr > 1.e-32 || cabr) {
                const double nT = 0.75 - 1.27 * lf;
                const double cT = -0.4 - 0.67 * lf;
                const double sT = log10(Pr) + cT;
                q = sT / (nT - 0.14 * sT);
            } else {
                q = 1. / 0.14;
            }
            wF = pow(10., lf / (1. + q * q));
        } else if (form == LSM_FALLOFF_SRI) {       // hpp:1283-1287, 1538-1544
            const double Fs = fp[3] * exp(-fp[4] / T) + exp(-T / fp[5]);
            const double lp = log10(Pr);
            wF = pow(Fs, 1. / (1. + lp * lp)) * fp[7];
            if (fp[6] != 0.) wF *= pow(T, fp[6]);
        }
        // chemically activated: k0 F/(1+Pr) (hpp:1613-1659, 1721-1726)
        if (cabr) kmod *= (1. / (1. + Pr)) * wF;
        else kmod *= kinf * (Pr / (1. + Pr)) * wF / kf;
    }
    // products of concentrations (StoichiometricMap.hpp:512-598)
    double pf = 1.;
    const bool sf = info & 8, sb = info & 16;     // non-elementary sides
    for (int q = M.fw_ptr[j]; q < M.fw_ptr[j + 1]; q++)
        pf *= sf ? order_pow(C(M.fw_sp[q]), M.fw_nu[q]) : conc_pow(C(M.fw_sp[q]), M.fw_nu[q]);
    double net = pf;
    if (info & LSM_RXN_REVERSIBLE) {
        // reaction dH/RT, dS/R (StoichiometricMap.hpp:916-1001) and 1/Kc (hpp:1215-1222)
        double dh = 0., ds = 0.;
        for (int q = M.rx_ptr[j]; q < M.rx_ptr[j + 1]; q++) {
            const int k = M.rx_sp[q];
            const double nu = M.rx_nu[q];
            dh += nu * H(k);
            ds += nu * S(k);
        }
        // explicit reverse Arrhenius: k_rev/k instead of 1/Kc (hpp:1242-1253, 1744-1750)
        const double *rv = M.rv + 3 * j;
        const double uKeq = (XK && (info & LSM_RXN_EXPLICIT_REV)) ? exp(rv[0] + rv[1] * logT - rv[2] * uT) / kf
                                                                : exp(-ds + dh - logPRT * M.rdn[j]);
        double pb = 1.;
        for (int q = M.bw_ptr[j]; q < M.bw_ptr[j + 1]; q++)
            pb *= sb ? order_pow(C(M.bw_sp[q]), M.bw_nu[q]) : conc_pow(C(M.bw_sp[q]), M.bw_nu[q]);
        net -= pb * uKeq;
    }
    return net * kmod;
}

// NASA-7 cp/R, h/RT, s/R of species k (ThermodynamicsMap_CHEMKIN.hpp:436-496)
__device__ __forceinline__ void nasa7(const DevMech &M, int k, double T, double T2, double T3, double T4,
                                      double uT, double logT, double &cp, double &h, double &s) {
    const double *sp = M.sp;
    const int P = M.ns_pad;
    const bool hi = T > sp[SP_TMID * P + k];
    const int rc = hi ? SP_CP_HT : SP_CP_LT, rh = hi ? SP_DH_HT : SP_DH_LT, rs = hi ? SP_DS_HT : SP_DS_LT;
    cp = sp[rc * P + k] + T * sp[(rc + 1) * P + k] + T2 * sp[(rc + 2) * P + k] + T3 * sp[(rc + 3) * P + k] +
         T4 * sp[(rc + 4) * P + k];
    h = sp[rh * P + k] + T * sp[(rh + 1) * P + k] + T2 * sp[(rh + 2) * P + k] + T3 * sp[(rh + 3) * P + k] +
        T4 * sp[(rh + 4) * P + k] + uT * sp[(rh + 5) * P + k];
    s = sp[rs * P + k] * logT + T * sp[(rs + 1) * P + k] + T2 * sp[(rs + 2) * P + k] +
        T3 * sp[(rs + 3) * P + k] + T4 * sp[(rs + 4) * P + k] + sp[(rs + 5) * P + k];
}

// ---------------------------------------------------------------------------
// Team context: one cell per wavefront (requires ne <= NMAX <= 64)
// ---------------------------------------------------------------------------
// Typed views of the mechanism image once it sits in LDS (layout: lsm_internal.hpp).
struct LdsMech {
    const double *sp, *P, *nuv, *tbe, *fop, *foe, *rvp, *pll, *chd;
    const uint4 *W;
    const unsigned *ent, *tbp, *tbs, *fol, *fos2, *plp, *chp;
    const int *fos;
    int SP;   // species stride (ns_pad)
    __device__ LdsMech(const unsigned char *b, const LdsImage &L, int ns_pad)
        : sp((const double *)(b + L.o_sp)), P((const double *)(b + L.o_P)), nuv((const double *)(b + L.o_nuv)),
          tbe((const double *)(b + L.o_tbe)), fop((const double *)(b + L.o_fop)), foe((const double *)(b + L.o_foe)),
          rvp((const double *)(b + L.o_rvp)), pll((const double *)(b + L.o_pll)), chd((const double *)(b + L.o_chd)),
          W((const uint4 *)(b + L.o_W)), ent((const unsigned *)(b + L.o_ent)), tbp((const unsigned *)(b + L.o_tbp)),
          tbs((const unsigned *)(b + L.o_tbs)), fol((const unsigned *)(b + L.o_fol)),
          fos2((const unsigned *)(b + L.o_fos2)), plp((const unsigned *)(b + L.o_plp)),
          chp((const unsigned *)(b + L.o_chp)), fos((const int *)(b + L.o_fos)),
          SP(ns_pad) {}
    __device__ __forceinline__ double nu(unsigned e) const {
        const unsigned c = e >> 12;
        return c < 16u ? 0.5 * (double)c : nuv[c - 16u];
    }
};

// One reaction (KineticsMap_CHEMKIN.hpp:1192-1386, 1683-1768), with the formation
// rates (StoichiometricMap.hpp:656-730) accumulated reaction-major into LDS.
// 1/Kc = exp(sum_k nu_k g_k - dn ln(Patm/RT)) with g = h/RT - s/R, which equals
// the reference's exp(-dS/R + dH/RT - dn ln(Patm/RT)).
// acc(k, v) adds v to the formation rate of species slot k.
template <bool XK, class Acc>
__device__ __forceinline__ void rxn_rate(const LdsMech &L, const double2 *scg, int j, double T, double logT, double uT,
                                         double logPRT, double Pk, double cTot, Acc &&acc) {
    const uint4 w = L.W[j];
    const double2 ab = ((const double2 *)L.P)[2 * j];
    const double2 ed = ((const double2 *)L.P)[2 * j + 1];
    const unsigned info = w.x;
    double kf;
    if (XK && (info & 128u)) {    // PLOG (beta field 0) or Chebyshev (1): slot in the lnA field
        const int s = (int)ab.x;
        double d;
        kf = ab.y == 0. ? exp(plog_lnk(L.pll + 4 * L.plp[s], (int)(L.plp[s + 1] - L.plp[s]), Pk, logT, uT, d))
                        : cheb_k(L.chd + L.chp[s], uT, Pk, d);
    } else {
        kf = exp(ab.x + ab.y * logT - ed.x * uT);
        if (info & 2u) kf = -kf;
    }
    double kmod = kf;
    const int tb = (int)((info >> 8) & 0xffu) - 1;
    if (tb >= 0) {
        double Me = cTot;
        for (unsigned q = L.tbp[tb]; q < L.tbp[tb + 1]; q++) Me += scg[L.tbs[q]].x * L.tbe[q];
        kmod *= Me;
    }
    const int fo = (int)((info >> 16) & 0xffu) - 1;
    if (fo >= 0) {
        const double *fp = L.fop + 8 * fo;
        const double kinf = exp(fp[0] + fp[1] * logT - fp[2] * uT);
        double Mc;
        if (L.fos[fo] < 0) {
            Mc = cTot;
            for (unsigned q = L.fol[fo]; q < L.fol[fo + 1]; q++) Mc += scg[L.fos2[q]].x * L.foe[q];
        } else {
            Mc = scg[L.fos[fo]].x + 1.e-16;
        }
        const double Pr = kf * Mc / kinf;
        const bool cabr = XK && (info & 64u);
        double wF = 1.;
        if (fp[7] > 0.) {    // Troe
            const double a = fp[3];
            double F = (1. - a) * exp(-T / fp[4]) + a * exp(-T / fp[5]);
            if (fp[6] != 0.) F += exp(-fp[6] / T);
            const double lf = (F < 1.e-300) ? -300. : log10(F);
            double q;
            if (Pr > 1.e-32 || cabr) {
                const double nT = 0.75 - 1.27 * lf;
                const double cT = -0.4 - 0.67 * lf;
                const double sT = log10(Pr) + cT;
                q = sT / (nT - 0.14 * sT);
            } else {
                q = 1. / 0.14;
            }
            wF = exp(LN10 * (lf / (1. + q * q)));    // 10^x
        } else if (fp[7] < 0.) {   // SRI: d [a exp(-b/T) + exp(-T/c)]^X T^e, X = 1/(1 + log10(Pr)^2)
            const double Fs = fp[3] * exp(-fp[4] * uT) + exp(-T / fp[5]);
            const double lp = log10(Pr);
            wF = exp(log(Fs) / (1. + lp * lp) + fp[6] * logT) * (-fp[7]);
        }
        if (cabr) kmod *= (1. / (1. + Pr)) * wF;    // chemically activated: k0 F/(1+Pr)
        else kmod *= kinf * (Pr / (1. + Pr)) * wF / kf;
    }
    if (info & 4u) {
        // fixed-format record: three reactant and three product slots
        const unsigned r0 = w.y & 0xffu, r1 = (w.y >> 8) & 0xffu, r2 = (w.y >> 16) & 0xffu;
        const unsigned p0 = w.z & 0xffu, p1 = (w.z >> 8) & 0xffu, p2 = (w.z >> 16) & 0xffu;
        const double2 a0 = scg[r0], a1 = scg[r1], a2 = scg[r2];
        const double2 b0 = scg[p0], b1 = scg[p1], b2 = scg[p2];
        double net = (a0.x * a1.x) * a2.x;
        if (info & 1u) {
            const double dg = (b0.y + b1.y + b2.y) - (a0.y + a1.y + a2.y);
            net -= ((b0.x * b1.x) * b2.x) * exp(dg - logPRT * ed.y);
        }
        const double r = net * kmod;
        acc(r0, -r);
        acc(r1, -r);
        acc(r2, -r);
        acc(p0, r);
        acc(p1, r);
        acc(p2, r);
        return;
    }
    // general stoichiometry (any orders, any number of species); explicit forward
    // orders, when given, follow the product entries
    const unsigned lst = w.w;
    const unsigned e0 = lst & 0xffffu, nre = (lst >> 16) & 0xffu, npr = lst >> 24;
    const unsigned nord = info >> 24;
    const bool sf = info & 8u, sb = info & 16u;
    const unsigned fb = nord ? e0 + nre + npr : e0, fn = nord ? nord : nre;
    double pf = 1.;
    for (unsigned q = 0; q < fn; q++) {
        const unsigned e = L.ent[fb + q];
        const double cq = scg[e & 0xfffu].x;
        pf *= sf ? order_pow(cq, L.nu(e)) : conc_pow(cq, L.nu(e));
    }
    double net = pf;
    if (info & 1u) {
        double dg = 0., pb = 1.;
        for (unsigned q = 0; q < nre; q++) {
            const unsigned e = L.ent[e0 + q];
            dg -= L.nu(e) * scg[e & 0xfffu].y;
        }
        for (unsigned q = 0; q < npr; q++) {
            const unsigned e = L.ent[e0 + nre + q];
            const double2 cg = scg[e & 0xfffu];
            const double nu = L.nu(e);
            dg += nu * cg.y;
            pb *= sb ? order_pow(cg.x, nu) : conc_pow(cg.x, nu);
        }
        if (XK && (info & 32u)) {    // explicit reverse Arrhenius (slot in the dn field): k_rev / k
            const double *rv = L.rvp + 4 * (int)ed.y;
            net -= pb * (exp(rv[0] + rv[1] * logT - rv[2] * uT) / kf);
        } else {
            net -= pb * exp(dg - logPRT * ed.y);
        }
    }
    const double r = net * kmod;
    for (unsigned q = 0; q < nre + npr; q++) {
        const unsigned e = L.ent[e0 + q];
        const double nu = L.nu(e);
        acc(e & 0xfffu, q < nre ? -nu * r : nu * r);
    }
}

// ---- analytic Jacobian ------------------------------------------------------
// One reaction's contributions to R_k, dR_k/dT at fixed c and dR_k/dc_i.  kmod is the effective rate constant
// (Arrhenius x M for third-body reactions, k_inf Pr/(1+Pr) F for fall-off),
// dkdM its derivative with respect to the collider concentration M.
// acc.R(k, r, drdT) adds to R_k and dR_k/dT, acc.D(i, k, v) to dR_k/dc_i, acc.U(k, v) to
// dR_k/dc_i for every species i (the part of a collider derivative dM/dc_i = 1 common to all
// species; the efficiencies eff_i - 1 go through acc.D); ns species,
// `dummy` the unused slot of fixed-format records (c = 1, g = 0), sH[k] = H_k/RT.
template <bool XK, class Acc>
__device__ __forceinline__ void rxn_jac(const LdsMech &L, const double2 *scg, const double *sH, int ns, unsigned dummy,
                                        int j, double T, double logT, double uT, double logPRT, double Pk,
                                        double cTot, Acc &&acc) {
    const uint4 w = L.W[j];
    const double2 ab = ((const double2 *)L.P)[2 * j];
    const double2 ed = ((const double2 *)L.P)[2 * j + 1];
    const unsigned info = w.x;
    double kf, dlnkf;
    if (XK && (info & 128u)) {    // PLOG or Chebyshev
        const int s = (int)ab.x;
        kf = ab.y == 0. ? exp(plog_lnk(L.pll + 4 * L.plp[s], (int)(L.plp[s + 1] - L.plp[s]), Pk, logT, uT, dlnkf))
                        : cheb_k(L.chd + L.chp[s], uT, Pk, dlnkf);
    } else {
        kf = exp(ab.x + ab.y * logT - ed.x * uT);
        if (info & 2u) kf = -kf;
        dlnkf = (ab.y + ed.x * uT) * uT;
    }
    double kmod = kf, dlnkmod = dlnkf, dkdM = 0.;
    int dense = 0, single = -1;
    const int tb = (int)((info >> 8) & 0xffu) - 1;
    const int fo = (int)((info >> 16) & 0xffu) - 1;
    if (tb >= 0) {
        double Me = cTot;
        for (unsigned q = L.tbp[tb]; q < L.tbp[tb + 1]; q++) Me += scg[L.tbs[q]].x * L.tbe[q];
        kmod = kf * Me;
        dkdM = kf;
        dense = 1;
    }
    if (fo >= 0) {
        const double *fp = L.fop + 8 * fo;
        const double kinf = exp(fp[0] + fp[1] * logT - fp[2] * uT);
        const double dlnkinf = (fp[1] + fp[2] * uT) * uT;
        double Mc;
        single = L.fos[fo];
        if (single < 0) {
            Mc = cTot;
            for (unsigned q = L.fol[fo]; q < L.fol[fo + 1]; q++) Mc += scg[L.fos2[q]].x * L.foe[q];
            dense = 2;
        } else {
            Mc = scg[single].x + 1.e-16;
        }
        const double Pr = kf * Mc / kinf;
        const bool cabr = XK && (info & 64u);
        double F = 1., dlgF_dlgPr = 0., dlnF_dT = 0.;
        if (fp[7] < 0.) {    // SRI: ln F = X ln Fs + e ln T + ln d, X = 1/(1 + u^2), u = log10 Pr
            const double eb = exp(-fp[4] * uT), ec = exp(-T / fp[5]);
            const double Fs = fp[3] * eb + ec;
            const double dFs = fp[3] * fp[4] * uT * uT * eb - ec / fp[5];
            const double u = log10(Pr);
            const double q = 1. + u * u;
            const double X = 1. / q;
            const double lnFs = log(Fs);
            F = exp(X * lnFs + fp[6] * logT) * (-fp[7]);
            dlgF_dlgPr = (Pr > 0.) ? -2. * u / (q * q) * lnFs / LN10 : 0.;   // d log10 F / d log10 Pr
            dlnF_dT = dlgF_dlgPr * (dlnkf - dlnkinf) + X * dFs / Fs + fp[6] * uT;
        }
        if (fp[7] > 0.) {    // Troe: log10 F = L / (1 + x^2), x = s / (n - 0.14 s), s = log10 Pr + c
            const double a = fp[3];
            const double e3 = exp(-T / fp[4]), e1 = exp(-T / fp[5]);
            double Fc = (1. - a) * e3 + a * e1;
            double dFc = -(1. - a) / fp[4] * e3 - a / fp[5] * e1;
            if (fp[6] != 0.) {
                const double e2 = exp(-fp[6] / T);
                Fc += e2;
                dFc += fp[6] * uT * uT * e2;
            }
            const double Lf = (Fc < 1.e-300) ? -300. : log10(Fc);
            const double dL = (Fc < 1.e-300) ? 0. : dFc / (Fc * LN10);
            double lgF, dlgF_dL;
            if (Pr > 1.e-32 || cabr) {
                const double n = 0.75 - 1.27 * Lf, cc = -0.4 - 0.67 * Lf;
                const double sT = log10(Pr) + cc;
                const double den = n - 0.14 * sT;
                const double x = sT / den;
                const double q = 1. + x * x;
                lgF = Lf / q;
                const double dlgF_dx = -Lf * 2. * x / (q * q);
                dlgF_dlgPr = dlgF_dx * n / (den * den);
                dlgF_dL = 1. / q + dlgF_dx * (1.27 * sT - 0.67 * n) / (den * den);
            } else {
                const double q = 1. + (1. / 0.14) * (1. / 0.14);
                lgF = Lf / q;
                dlgF_dL = 1. / q;
            }
            F = exp(LN10 * lgF);
            dlnF_dT = dlgF_dlgPr * (dlnkf - dlnkinf) + LN10 * dlgF_dL * dL;
        }
        if (cabr) {    // k0 F/(1+Pr): d/dM = k0 F/(1+Pr) (dlog F/dlog Pr / M - (k0/k_inf)/(1+Pr))
            kmod = kf * F / (1. + Pr);
            dkdM = kmod * (dlgF_dlgPr / Mc - (kf / kinf) / (1. + Pr));
            dlnkmod = dlnkf + dlnF_dT - (Pr / (1. + Pr)) * (dlnkf - dlnkinf);
        } else {
            kmod = kinf * (Pr / (1. + Pr)) * F;
            dkdM = (kf * F / (1. + Pr)) * (1. / (1. + Pr) + dlgF_dlgPr);   // d kmod / d M
            dlnkmod = dlnkinf + (dlnkf - dlnkinf) / (1. + Pr) + dlnF_dT;
        }
    }
    if (info & 4u) {
        const unsigned sl[6] = {w.y & 0xffu, (w.y >> 8) & 0xffu, (w.y >> 16) & 0xffu,
                                w.z & 0xffu, (w.z >> 8) & 0xffu, (w.z >> 16) & 0xffu};
        double cv[6];
#pragma unroll
        for (int e = 0; e < 6; e++) cv[e] = scg[sl[e]].x;
        const double pf = (cv[0] * cv[1]) * cv[2];
        double g[6];
        g[0] = kmod * (cv[1] * cv[2]);
        g[1] = kmod * (cv[0] * cv[2]);
        g[2] = kmod * (cv[0] * cv[1]);
        double uK = 0., pb = 0., dh = 0.;
        g[3] = g[4] = g[5] = 0.;
        if (info & 1u) {
            const double dg = (scg[sl[3]].y + scg[sl[4]].y + scg[sl[5]].y) - (scg[sl[0]].y + scg[sl[1]].y + scg[sl[2]].y);
            uK = exp(dg - logPRT * ed.y);
            pb = (cv[3] * cv[4]) * cv[5];
            dh = (sH[sl[3]] + sH[sl[4]] + sH[sl[5]]) - (sH[sl[0]] + sH[sl[1]] + sH[sl[2]]);
            const double kb = -kmod * uK;
            g[3] = kb * (cv[4] * cv[5]);
            g[4] = kb * (cv[3] * cv[5]);
            g[5] = kb * (cv[3] * cv[4]);
        }
        const double net = pf - pb * uK;
        const double r = kmod * net;
        const double drdT = kmod * (dlnkmod * net - pb * uK * (ed.y - dh) * uT);
#pragma unroll
        for (int o = 0; o < 6; o++) {
            acc.R(sl[o], o < 3 ? -r : r, o < 3 ? -drdT : drdT);
        }
#pragma unroll
        for (int d = 0; d < 6; d++) {
            if (sl[d] == dummy) continue;
#pragma unroll
            for (int o = 0; o < 6; o++)
                if (sl[o] != dummy) acc.D(sl[d], sl[o], o < 3 ? -g[d] : g[d]);
        }
        if (dense) {
            const double qv = dkdM * net;
#pragma unroll
            for (int o = 0; o < 6; o++)
                if (sl[o] != dummy) acc.U(sl[o], o < 3 ? -qv : qv);
            const unsigned q0 = dense == 1 ? L.tbp[tb] : L.fol[fo], q1 = dense == 1 ? L.tbp[tb + 1] : L.fol[fo + 1];
            for (unsigned q = q0; q < q1; q++) {
                const unsigned i = dense == 1 ? L.tbs[q] : L.fos2[q];
                const double e1 = qv * (dense == 1 ? L.tbe[q] : L.foe[q]);
#pragma unroll
                for (int o = 0; o < 6; o++)
                    if (sl[o] != dummy) acc.D(i, sl[o], o < 3 ? -e1 : e1);
            }
        } else if (single >= 0) {
            const double qv = dkdM * net;
#pragma unroll
            for (int o = 0; o < 6; o++)
                if (sl[o] != dummy) acc.D(single, sl[o], o < 3 ? -qv : qv);
        }
        return;
    }
    // general stoichiometry: d(prod f(c))/dc_i = f'(c_i) prod_{others} f(c), f = c^nu or,
    // on a non-elementary side, order_pow; explicit forward orders follow the products
    const unsigned lst = w.w;
    const unsigned e0 = lst & 0xffffu, nre = (lst >> 16) & 0xffu, npr = lst >> 24;
    const unsigned nord = info >> 24;
    const bool sf = info & 8u, sb = info & 16u;
    const unsigned fb = nord ? e0 + nre + npr : e0, fn = nord ? nord : nre;
    double pf = 1., pb = 1., dg = 0., dh = 0., uK = 0., dlnuK = 0.;
    for (unsigned q = 0; q < fn; q++) {
        const unsigned e = L.ent[fb + q];
        const double cq = scg[e & 0xfffu].x;
        pf *= sf ? order_pow(cq, L.nu(e)) : conc_pow(cq, L.nu(e));
    }
    if (info & 1u) {
        for (unsigned q = 0; q < nre + npr; q++) {
            const unsigned e = L.ent[e0 + q];
            const double nu = q < nre ? -L.nu(e) : L.nu(e);
            dg += nu * scg[e & 0xfffu].y;
            dh += nu * sH[e & 0xfffu];
            if (q >= nre) pb *= sb ? order_pow(scg[e & 0xfffu].x, nu) : conc_pow(scg[e & 0xfffu].x, nu);
        }
        if (XK && (info & 32u)) {    // explicit reverse: uK = k_rev / k
            const double *rv = L.rvp + 4 * (int)ed.y;
            uK = exp(rv[0] + rv[1] * logT - rv[2] * uT) / kf;
            dlnuK = (rv[1] + rv[2] * uT) * uT - dlnkf;
        } else {
            uK = exp(dg - logPRT * ed.y);
            dlnuK = (ed.y - dh) * uT;
        }
    }
    const double net = pf - pb * uK;
    const double r = kmod * net;
    const double drdT = XK ? kmod * (dlnkmod * net - pb * uK * dlnuK)
                           : kmod * (dlnkmod * net - pb * uK * (ed.y - dh) * uT);
    for (unsigned q = 0; q < nre + npr; q++) {
        const unsigned e = L.ent[e0 + q];
        const double nu = L.nu(e);
        acc.R(e & 0xfffu, q < nre ? -nu * r : nu * r, q < nre ? -nu * drdT : nu * drdT);
    }
    for (unsigned d = 0; d < fn + npr; d++) {
        const bool fwd = d < fn;
        if (!fwd && !(info & 1u)) break;
        const unsigned base = fwd ? fb : e0 + nre;          // entries of this side
        const unsigned cnt = fwd ? fn : npr, dd = fwd ? d : d - fn;
        const bool sm = fwd ? sf : sb;
        const unsigned ed_ = L.ent[base + dd];
        const unsigned id = ed_ & 0xfffu;
        const double nud = L.nu(ed_), cd = scg[id].x;
        // d(c^nu)/dc; at c = 0 a fractional order has no finite slope: 0 is used
        double der = sm ? order_pow_der(cd, nud)
                        : (nud == 1. ? 1. : ((cd > 0. || nud > 1.) ? nud * conc_pow(cd, nud - 1.) : 0.));
        for (unsigned q = 0; q < cnt; q++)
            if (q != dd) {
                const unsigned eq = L.ent[base + q];
                const double cq = scg[eq & 0xfffu].x;
                der *= sm ? order_pow(cq, L.nu(eq)) : conc_pow(cq, L.nu(eq));
            }
        const double gv = fwd ? kmod * der : -kmod * uK * der;
        for (unsigned o = 0; o < nre + npr; o++) {
            const unsigned eo = L.ent[e0 + o];
            const double nu = L.nu(eo);
            acc.D(id, eo & 0xfffu, o < nre ? -nu * gv : nu * gv);
        }
    }
    if (dense || single >= 0) {
        const double qv = dkdM * net;
        for (unsigned o = 0; o < nre + npr; o++) {
            const unsigned eo = L.ent[e0 + o];
            const double sg = o < nre ? -L.nu(eo) : L.nu(eo);
            const unsigned ko = eo & 0xfffu;
            if (single >= 0) {
                acc.D(single, ko, sg * qv);
                continue;
            }
            acc.U(ko, sg * qv);
            const unsigned q0 = dense == 1 ? L.tbp[tb] : L.fol[fo], q1 = dense == 1 ? L.tbp[tb + 1] : L.fol[fo + 1];
            for (unsigned q = q0; q < q1; q++) {
                const unsigned i = dense == 1 ? L.tbs[q] : L.fos2[q];
                acc.D(i, ko, sg * qv * (dense == 1 ? L.tbe[q] : L.foe[q]));
            }
        }
    }
}

template <int NMAX, bool XK>
struct TeamCtx {
    double V[NVEC][1];
    double G[NMAX];  // this lane's row of the LU factors of the Newton matrix: multipliers in the
                     // columns eliminated before the row became the pivot (k < ps), U from ps on.
                     // Kept in registers from one factorisation to the next.
    int ps;          // elimination step at which this lane's row became the pivot
    int plv;         // lane k holds the row (lane) chosen as pivot at step k
    double dinv;     // 1 / pivot of this lane's row
    double *Jl;                       // per-wave LDS: Jacobian [ne][65], lane k holds row k (df_k/dy_j at
                                      // [j*65 + k]); during an analytic Jacobian first dR_k/dc_j
    double2 *scg;                     // per-wave LDS: (c_k, g_k = h_k/RT - s_k/R) per species slot
    double *sR;                       // per-wave LDS: formation-rate accumulators
    double *sDT;                      // per-wave LDS: dR_k/dT accumulators (analytic Jacobian)
    double *sH;                       // per-wave LDS: h_k = H_k/RT (analytic Jacobian)
    double *sU;                       // per-wave LDS: dR_k/dc_i common to every i (collider terms)
    LdsMech L;
    const KParams &K;
    int lane, ne, ns, nr;
    int s0;                           // first reaction of the paired loop (LdsImage::simple0)
    double qr;
    double rho0;                      // constant-volume reactor: density of the current cell
#ifdef LSM_PHASE_TIMERS
    unsigned long long tp[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    __device__ __forceinline__ void add_time(int i, unsigned long long d) { tp[i] += d; }
    int in_jac = 0;
#endif

    __device__ TeamCtx(const DevMech &m, const KParams &k, const unsigned char *img, double *wave_lds)
        : L(img, m.img, m.ns_pad), K(k), lane(threadIdx.x & 63), ne(m.ne), ns(m.ns), nr(m.nr), s0(m.img.simple0),
          qr(0.), rho0(0.) {
        scg = (double2 *)wave_lds;
        sR = wave_lds + 128;
        sDT = wave_lds + 192;
        sH = wave_lds + 256;
        sU = wave_lds + 320;
        Jl = wave_lds + TEAM_WAVE_FIXED;
    }

    template <class F>
    __device__ __forceinline__ void each(F f) { f(0); }
    __device__ __forceinline__ double &x(int vid, int e) { return V[vid][e]; }
    __device__ __forceinline__ double ymin(int) const {
        return lane < ns ? 0. : (lane == ns ? 200. : NEG_INF);
    }
    __device__ __forceinline__ double ymax(int) const {
        return lane < ns ? 1. : (lane == ns ? 6000. : POS_INF);
    }
    __device__ __forceinline__ double reduce(double v) { return wave_sum(v); }
    __device__ __forceinline__ int count(bool b) { return __popcll(__ballot(b)); }

    __device__ __forceinline__ void reaction(int j, double T, double logT, double uT, double logPRT, double Pk,
                                            double cTot) {
        rxn_rate<XK>(L, scg, j, T, logT, uT, logPRT, Pk, cTot, [&](unsigned k, double v) { atomicAdd(racc(k), v); });
    }

    // Accumulator of species slot k in the reaction loop.  The unused slots of fixed-format records
    // all name the dummy species; their (zero-rate-weighted) adds go to a slot of each lane's own
    // (sU, free during an RHS) so that they do not serialise on one LDS address.
    __device__ __forceinline__ double *racc(unsigned k) const {
        return k == (unsigned)LSM_DUMMY_SP ? sU + lane : sR + k;
    }

    // Two fixed-format records without collider (reactions j and j + 64, the second absent past nr),
    // their loads, exponentials and products interleaved: the same arithmetic as rxn_rate's
    // fixed-format path (KineticsMap_CHEMKIN.hpp:1198-1222, 1734-1767).
    __device__ __forceinline__ void simple_pair(int j, double logT, double uT, double logPRT) {
        const int j2 = j + 64;
        const bool has2 = j2 < nr;
        const int jb = has2 ? j2 : j;
        const uint4 w1 = L.W[j], w2 = L.W[jb];
        const double2 ab1 = ((const double2 *)L.P)[2 * j], ed1 = ((const double2 *)L.P)[2 * j + 1];
        const double2 ab2 = ((const double2 *)L.P)[2 * jb], ed2 = ((const double2 *)L.P)[2 * jb + 1];
        const unsigned r10 = w1.y & 0xffu, r11 = (w1.y >> 8) & 0xffu, r12 = (w1.y >> 16) & 0xffu;
        const unsigned p10 = w1.z & 0xffu, p11 = (w1.z >> 8) & 0xffu, p12 = (w1.z >> 16) & 0xffu;
        const unsigned r20 = w2.y & 0xffu, r21 = (w2.y >> 8) & 0xffu, r22 = (w2.y >> 16) & 0xffu;
        const unsigned p20 = w2.z & 0xffu, p21 = (w2.z >> 8) & 0xffu, p22 = (w2.z >> 16) & 0xffu;
        const double2 a10 = scg[r10], a11 = scg[r11], a12 = scg[r12], b10 = scg[p10], b11 = scg[p11], b12 = scg[p12];
        const double2 a20 = scg[r20], a21 = scg[r21], a22 = scg[r22], b20 = scg[p20], b21 = scg[p21], b22 = scg[p22];
        double kf1 = exp(ab1.x + ab1.y * logT - ed1.x * uT);
        double kf2 = exp(ab2.x + ab2.y * logT - ed2.x * uT);
        kf1 = (w1.x & 2u) ? -kf1 : kf1;
        kf2 = (w2.x & 2u) ? -kf2 : kf2;
        const double dg1 = (b10.y + b11.y + b12.y) - (a10.y + a11.y + a12.y);
        const double dg2 = (b20.y + b21.y + b22.y) - (a20.y + a21.y + a22.y);
        const double uk1 = exp(dg1 - logPRT * ed1.y), uk2 = exp(dg2 - logPRT * ed2.y);
        double net1 = (a10.x * a11.x) * a12.x, net2 = (a20.x * a21.x) * a22.x;
        const double rb1 = ((b10.x * b11.x) * b12.x) * uk1, rb2 = ((b20.x * b21.x) * b22.x) * uk2;
        net1 = (w1.x & 1u) ? net1 - rb1 : net1;
        net2 = (w2.x & 1u) ? net2 - rb2 : net2;
        const double r1 = net1 * kf1;
        const double r2 = has2 ? net2 * kf2 : 0.;
        atomicAdd(racc(r10), -r1); atomicAdd(racc(r11), -r1); atomicAdd(racc(r12), -r1);
        atomicAdd(racc(p10), r1);  atomicAdd(racc(p11), r1);  atomicAdd(racc(p12), r1);
        atomicAdd(racc(r20), -r2); atomicAdd(racc(r21), -r2); atomicAdd(racc(r22), -r2);
        atomicAdd(racc(p20), r2);  atomicAdd(racc(p21), r2);  atomicAdd(racc(p22), r2);
    }

    __device__ __forceinline__ void rhs(int src, int dst) {
        LSM_TIC(t0);
        rhs_impl(src, dst);
#ifdef LSM_PHASE_TIMERS
        if (!in_jac) LSM_TOC(0, t0);
#endif
    }

    // BatchReactorHomogeneousConstantPressure::Equations (:175-247), or with K.const_volume
    // BatchReactorHomogeneousConstantVolume::Equations (BatchReactorHomogeneousConstantVolume.H:150-199)
    __device__ __forceinline__ void rhs_impl(int src, int dst) {
        LSM_TIC(t_rhs0);
        const double T = readlane_d(V[src][0], ns);
        const int k = lane;
        const bool isp = k < ns;
        const double Yk = isp ? V[src][0] : 0.;
        const int P = L.SP;
        const double mwk = isp ? L.sp[15 * P + k] : 1.;
        const double MW = 1. / wave_sum(Yk / mwk);
        const double xk = (MW * Yk) / mwk;
        const double cTotP = K.const_volume ? rho0 / MW : K.pressure / RGAS / T;
        const double ck = cTotP * xk;
        const double rho = K.const_volume ? rho0 : cTotP * MW;
        const double T2 = T * T, T3 = T2 * T, T4 = T3 * T, uT = 1. / T, logT = log(T);
        double cp = 0., h = 0.;
        if (isp) {
            // NASA-7 (ThermodynamicsMap_CHEMKIN.hpp:263-343, 436-496), high set iff T > T_mid
            const double *a = L.sp + ((T > L.sp[14 * P + k]) ? 0 : 7 * P) + k;
            const double a0 = a[0], a1 = a[P], a2 = a[2 * P], a3 = a[3 * P], a4 = a[4 * P];
            cp = a0 + T * a1 + T2 * a2 + T3 * a3 + T4 * a4;
            h = a0 + T * (0.50 * a1) + T2 * ((1. / 3.) * a2) + T3 * (0.25 * a3) + T4 * (0.20 * a4) + uT * a[5 * P];
            const double s = a0 * logT + T * a1 + T2 * (0.50 * a2) + T3 * ((1. / 3.) * a3) + T4 * (0.25 * a4) +
                             a[6 * P];
            scg[k] = make_double2(ck, h - s);
        }
        if (k == LSM_DUMMY_SP) scg[k] = make_double2(1., 0.);
        sR[k] = 0.;
        double cTot = ck, cpx = isp ? cp * xk : 0.;
        wave_sum2(cTot, cpx);
        wave_sync();
        LSM_TOC(7, t_rhs0);
        LSM_TIC(t_rhs1);
        const double logPRT = log(101325. / RGAS / T);
        // the kinetics' pressure: P0, or cTot R T at constant volume
        // (BatchReactorHomogeneousConstantVolume.H:166-176)
        const double Pk = K.const_volume ? cTotP * RGAS * T : K.pressure;
        for (int j = lane; j < s0; j += 64) reaction(j, T, logT, uT, logPRT, Pk, cTot);
        for (int j = s0 + lane; j < nr; j += 128) simple_pair(j, logT, uT, logPRT);
        wave_sync();
        LSM_TOC(8, t_rhs1);
        LSM_TIC(t_rhs2);
        const double Rk = isp ? sR[k] : 0.;
        double dy = isp ? (mwk * Rk) / rho : 0.;
        if (K.energy && K.const_volume) {
            // dT/dt = (QR + R T sum_k R_k) / (rho0 cv), cv = (CpMixMolar - R) / MW
            double hs = isp ? Rk * h : 0., rs = Rk;
            wave_sum2(hs, rs);
            qr = -hs * RGAS * T;
            const double CvMixMass = (cpx * RGAS - RGAS) / MW;
            if (k == ns) dy = (qr + RGAS * T * rs) / (rho * CvMixMass);
        } else if (K.energy) {
            const double hs = wave_sum(isp ? Rk * h : 0.);
            qr = -hs * RGAS * T;
            const double CpMixMass = (cpx * RGAS) / MW;
            if (k == ns) dy = qr / (rho * CpMixMass);
        }
        wave_sync();
        V[dst][0] = dy;
        LSM_TOC(9, t_rhs2);
    }

    // Constant-volume reactor: density of the cell (rhoCells, chemistry_DI.H:173), or
    // the ideal-gas density of the initial state y0 (lane value) at the reactor pressure.
    __device__ __forceinline__ void set_density(double y0, double T0, const double *rho_cells, int64_t cell) {
        const double mwk = lane < ns ? L.sp[15 * L.SP + lane] : 1.;
        const double MW0 = 1. / wave_sum(lane < ns ? y0 / mwk : 0.);
        rho0 = rho_cells ? rho_cells[cell] : K.pressure * MW0 / (RGAS * T0);
    }

    // Strang "compact" sources of the final state y (chemistry_DI.H:236-282):
    // RR_k = rho (Y_k - Y_k,in) / dt, RT = rho cp (T - T_in) / dt, rho and cp of the final
    // state at the reactor pressure; zero for dt <= 1e-14, RT zero without energy equation.
    __device__ __forceinline__ double compact_sources(double y, double yin, double Tin, double dtc) {
        const bool isp = lane < ns;
        const double Tf = readlane_d(y, ns);
        const int P = L.SP;
        const double mwk = isp ? L.sp[15 * P + lane] : 1.;
        double cp = 0.;
        if (isp) {
            const double *a = L.sp + ((Tf > L.sp[14 * P + lane]) ? 0 : 7 * P) + lane;
            cp = a[0] + Tf * (a[P] + Tf * (a[2 * P] + Tf * (a[3 * P] + Tf * a[4 * P])));
        }
        double u = isp ? y / mwk : 0., cy = isp ? cp * (y / mwk) : 0.;
        wave_sum2(u, cy);
        const double rho = K.pressure * (1. / u) / (RGAS * Tf);
        const bool ok = dtc > 1.e-14;
        if (isp) return ok ? rho * (y - yin) / dtc : 0.;
        return (ok && K.energy) ? rho * (cy * RGAS) * (Tf - Tin) / dtc : 0.;
    }

    // Accumulators of the analytic Jacobian: R and dR/dT, dR_k/dc_i into the Jacobian's own
    // LDS ([i*65 + k]), the collider terms common to every column into sU.
    struct JacAcc {
        double *sR, *sDT, *Dm, *sU;
        __device__ __forceinline__ void R(unsigned k, double r, double drdT) const {
            if (k != (unsigned)LSM_DUMMY_SP) {     // the dummy species of fixed-format records
                atomicAdd(&sR[k], r);
                atomicAdd(&sDT[k], drdT);
            }
        }
        __device__ __forceinline__ void D(unsigned i, unsigned k, double v) const { atomicAdd(&Dm[i * 65 + k], v); }
        __device__ __forceinline__ void U(unsigned k, double v) const { atomicAdd(&sU[k], v); }
    };

    // Analytic Jacobian of the constant-pressure reactor at the state in VB:
    //   f_k = MW_k R_k / rho,  f_T = -sum_k H_k R_k / (rho cp)
    // with D = dR/dc (T fixed), S_k = sum_i D_ki c_i:
    //   df_k/dY_j = (MW_k / MW_j) D_kj + (MW / MW_j) (f_k - MW_k S_k / rho)
    //   df_k/dT   = MW_k / rho (dR_k/dT - S_k / T) + f_k / T
    //   df_T/dY_j = [-(RT / MW_j)(rho sum_k h_k D_kj - MW sum_k h_k S_k) - f_T (rho / MW_j)(Cp_j - MW cp)] / (rho cp)
    //   df_T/dT   = [-sum_k (Cp_k R_k + H_k (dR_k/dT - S_k / T)) - f_T (-rho cp / T + rho sum_k Y_k Cp'_k / MW_k)] / (rho cp)
    // (MW mixture molecular weight, cp mass heat capacity, h_k = H_k / RT).  D is gathered in
    // the Jacobian's LDS (the collider terms common to every column apart, in sU), its sums S and
    // sum_k h_k D_kj are taken, and each lane then turns its own row of D into its row of J.
    __device__ __forceinline__ void jacobian_analytic_impl() {
        const double T = readlane_d(V[VB_][0], ns);
        const int k = lane;
        const bool isp = k < ns;
        const double Yk = isp ? V[VB_][0] : 0.;
        const int P = L.SP;
        const double mwk = isp ? L.sp[15 * P + k] : 1.;
        const double MW = 1. / wave_sum(Yk / mwk);
        const double xk = (MW * Yk) / mwk;
        const double cTotP = K.pressure / RGAS / T;
        const double ck = cTotP * xk;
        const double rho = cTotP * MW;
        const double T2 = T * T, T3 = T2 * T, T4 = T3 * T, uT = 1. / T, logT = log(T);
        double cp = 0., h = 0., dcp = 0.;
        if (isp) {
            const double *a = L.sp + ((T > L.sp[14 * P + k]) ? 0 : 7 * P) + k;
            const double a0 = a[0], a1 = a[P], a2 = a[2 * P], a3 = a[3 * P], a4 = a[4 * P];
            cp = a0 + T * a1 + T2 * a2 + T3 * a3 + T4 * a4;
            dcp = a1 + T * (2. * a2) + T2 * (3. * a3) + T3 * (4. * a4);
            h = a0 + T * (0.50 * a1) + T2 * ((1. / 3.) * a2) + T3 * (0.25 * a3) + T4 * (0.20 * a4) + uT * a[5 * P];
            const double s = a0 * logT + T * a1 + T2 * (0.50 * a2) + T3 * ((1. / 3.) * a3) + T4 * (0.25 * a4) +
                             a[6 * P];
            scg[k] = make_double2(ck, h - s);
            sH[k] = h;
        }
        if (k == LSM_DUMMY_SP) { scg[k] = make_double2(1., 0.); sH[k] = 0.; }
        sR[k] = 0.;
        sDT[k] = 0.;
        sU[k] = 0.;
        for (int i = 0; i < ne; i++) Jl[i * 65 + k] = 0.;
        double cTot = ck, cpx = isp ? cp * xk : 0.;
        wave_sum2(cTot, cpx);
        wave_sync();
        const double logPRT = log(101325. / RGAS / T);
        const JacAcc acc{sR, sDT, Jl, sU};
        for (int j = lane; j < nr; j += 64)
            rxn_jac<XK>(L, scg, sH, ns, LSM_DUMMY_SP, j, T, logT, uT, logPRT, K.pressure, cTot, acc);
        wave_sync();
        const double Rk = isp ? sR[k] : 0.;
        const double dRdT = isp ? sDT[k] : 0.;
        const double Uk = isp ? sU[k] : 0.;
        double S = 0., A = 0.;
        if (isp) {
            for (int i = 0; i < ns; i++) S += (Jl[i * 65 + k] + Uk) * scg[i].x;       // own row of D
            for (int i = 0; i < ns; i++) A += sH[i] * (Jl[k * 65 + i] + sU[i]);      // sum_i h_i dR_i/dc_k
        }
        wave_sync();                                                                  // D read before rewritten
        const double fk = isp ? (mwk * Rk) / rho : 0.;
        if (isp) {
            const double g = fk - mwk * S / rho;
            for (int j = 0; j < ns; j++) {
                const double mwj = L.sp[15 * P + j];
                Jl[j * 65 + k] = (mwk / mwj) * (Jl[j * 65 + k] + Uk) + (MW / mwj) * g;
            }
            Jl[ns * 65 + k] = (mwk / rho) * (dRdT - S * uT) + fk * uT;
        } else if (k > ns) {
            for (int j = 0; j < ne; j++) Jl[j * 65 + k] = 0.;
        }
        // temperature row (zero without the energy equation)
        const double cpm = (cpx * RGAS) / MW;            // mass cp
        const double Phi = rho * cpm;
        double hs = isp ? h * Rk : 0., hS = isp ? h * S : 0.;
        wave_sum2(hs, hS);
        const double QR = -hs * RGAS * T;
        const double fT = K.energy ? QR / Phi : 0.;
        double dQ = isp ? -(cp * RGAS * Rk + h * RGAS * T * (dRdT - S * uT)) : 0.;
        double dP = isp ? Yk * dcp * RGAS / mwk : 0.;
        wave_sum2(dQ, dP);
        if (K.energy) {
            if (isp) {
                const double dQdY = -(RGAS * T / mwk) * (rho * A - MW * hS);
                const double dPdY = (rho / mwk) * (cp * RGAS - MW * cpm);
                Jl[k * 65 + ns] = (dQdY - fT * dPdY) / Phi;
            }
            if (k == ns) Jl[ns * 65 + ns] = (dQ - fT * (-Phi * uT + rho * dP)) / Phi;
        } else {
            if (isp) Jl[k * 65 + ns] = 0.;
            if (k == ns) Jl[ns * 65 + ns] = 0.;
        }
        wave_sync();
    }

    // KernelDense::NumericalJacobian, max-constraint branch (KernelDense.hpp:160-183)
    __device__ __forceinline__ void jacobian_impl(double h) {
        const double ZERO_DER = sqrt(TINY_F), ETA2 = sqrt(EPS_D), BETA = 1.e+3 * EPS_D;
        double fw = V[F_][0] * V[OOE_][0];
        double hf = BETA * fabs(h) * sqrt(wave_sum(fw * fw) / (double)ne) * (double)ne;
        if (hf < 1.e-10) hf = 1.;
        for (int j = 0; j < ne; j++) {
            const double yh = readlane_d(V[VB_][0], j);
            const double ej = readlane_d(V[OOE_][0], j);
            const double a = fabs(yh), b = fabs(1. / ej);
            double hJ = ETA2 * (a > b ? a : b);
            hJ = fmax(hJ, hf / ej);
            hJ = fmax(hJ, ZERO_DER);
            const double ymx = j < ns ? 1. : 6000.;
            if (yh + hJ > ymx) hJ = -hJ;
            if (lane == j) V[VB_][0] = yh + hJ;
            rhs(VB_, AUX_);
            if (lane == j) V[VB_][0] = yh;
            const double hInv = 1. / hJ;
            Jl[j * 65 + lane] = hInv * (V[AUX_][0] - V[F_][0]);
        }
    }

    // Row of the largest |g| among the eligible lanes (the lowest lane on a tie; the first
    // eligible lane when every eligible |g| is 0), as a uniform value.
    __device__ __forceinline__ int pivot_lane(double g, bool elig) {
        const double a = elig ? fabs(g) : 0.;
        const double m = wave_max_nonneg(a);
        const unsigned long long hit = __ballot(elig && a == m);
        return __builtin_amdgcn_readfirstlane(__ffsll(hit) - 1);
    }

    // G = I - hr0 J and row-pivoted LU (KernelDense.hpp:196-221), factorised in
    // registers (lane i holds row i, the pivot row is broadcast with v_readlane)
    // and kept there for the solves.  Rows are not moved: lane i records the
    // step ps at which it was chosen; the largest |G(i,k)| among rows not yet
    // chosen wins (the lowest lane on a tie).
    __device__ __forceinline__ void factor_impl(double hr0) {
        // Opaque copies of lane and ne: without them the compiler hoists the NMAX
        // identity entries and the NMAX "j < ne" masks out of the persistent cell
        // loop and keeps them live (in AGPRs / spilled SGPRs) for the whole kernel.
        int lane = this->lane, ne = this->ne;
        asm volatile("" : "+v"(lane));
        asm volatile("" : "+s"(ne));
#pragma unroll
        for (int j = 0; j < NMAX; j++) {
            const double jij = (j < ne) ? Jl[j * 65 + lane] : 0.;
            G[j] = jij * -hr0 + ((lane == j) ? 1. : 0.);
            if ((j & 7) == 7) __builtin_amdgcn_sched_barrier(0);
        }
        ps = (lane < ne) ? -1 : 1 << 20;
        dinv = 1.;
        // The pivot of step k is the largest |G(i,k)| among rows not yet chosen, the lowest lane
        // on a tie (or the first eligible row when the column is zero).  Each step's search runs
        // as soon as its column has been updated, overlapping the rest of the previous update.
        int p = pivot_lane(G[0], ps < 0);
#pragma unroll
        for (int k = 0; k < NMAX; k++) {
            if (k < ne) {
                const double piv = readlane_d(G[k], p);
                const double inv = 1. / piv;
                if (lane == p) { ps = k; dinv = inv; }
                double l = 0.;
                if (ps < 0 && piv != 0.) {
                    l = G[k] * inv;
                    G[k] = l;
                }
                int pn = 0;
                if (k + 1 < NMAX) {
                    G[k + 1] = fma(-l, readlane_d(G[k + 1], p), G[k + 1]);
                    pn = pivot_lane(G[k + 1], ps < 0);
                }
                // four columns at a time, their pivot-row elements read first: the barrier keeps
                // the compiler from interleaving read and update column by column through a
                // single scalar register pair (groups of 8 need more scalar registers than are free)
#pragma unroll
                for (int j = k + 2; j < NMAX; j += 4) {
                    double u[4];
#pragma unroll
                    for (int t = 0; t < 4; t++)
                        if (j + t < NMAX) u[t] = readlane_d(G[j + t], p);
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int t = 0; t < 4; t++)
                        if (j + t < NMAX) G[j + t] = fma(-l, u[t], G[j + t]);
                }
                p = pn;
            }
        }
        // pivot order: lane k learns which row was chosen at step k (sR is free here)
        int *pl = (int *)sR;
        if (lane < ne) pl[ps] = lane;
        wave_sync();
        plv = (lane < ne) ? pl[lane] : 0;
        wave_sync();
    }

    // Forward / backward substitution in pivot order with the factors in registers, in groups of
    // 8 steps with no branch inside a group: the group's masked factors and pivot lanes are formed
    // first, off the dependent chain b -> broadcast -> fma.  Steps past ne inside the last group
    // change nothing: rows of the system have zero factors there, and the padding rows (lanes
    // >= ne) are never read.  Masks are products, not selects, which the compiler would sink into
    // branches.
    __device__ __forceinline__ void solve_impl(int vid) {
        // Opaque per-solve copies of the pivot data: otherwise the compiler hoists the masked
        // factors, pivot lanes and inverse pivots of every step out of the Newton and step loops
        // (they only change at a factorisation) and keeps them all live, far beyond the
        // register budget.
        int ne = this->ne, ps_l = ps, plv_l = plv;
        double dinv_l = dinv;
        asm volatile("" : "+s"(ne));
        asm volatile("" : "+v"(ps_l), "+v"(plv_l), "+v"(dinv_l));
        double b = V[vid][0];
#pragma unroll
        for (int k0 = 0; k0 < NMAX; k0 += 8) {
            if (k0 < ne) {
                double lk[8];
                int pk[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    lk[u] = G[k0 + u] * ((ps_l > k0 + u) ? 1. : 0.);
                    pk[u] = __builtin_amdgcn_readlane(plv_l, k0 + u);
                }
#pragma unroll
                for (int u = 0; u < 8; u++) b = fma(-lk[u], readlane_d(b, pk[u]), b);
            }
        }
        double xr = 0.;
#pragma unroll
        for (int k0 = NMAX - 8; k0 >= 0; k0 -= 8) {
            if (k0 < ne) {
                double uk[8], dk[8];
                int pk[8];
#pragma unroll
                for (int u = 7; u >= 0; u--) {
                    uk[u] = G[k0 + u] * ((ps_l < k0 + u) ? 1. : 0.);
                    pk[u] = __builtin_amdgcn_readlane(plv_l, k0 + u);
                    dk[u] = readlane_d(dinv_l, pk[u]);
                }
#pragma unroll
                for (int u = 7; u >= 0; u--) {
                    const double xk = readlane_d(b, pk[u]) * dk[u];
                    b = fma(-uk[u], xk, b);
                    if (lane == k0 + u) xr = xk;
                }
            }
        }
        V[vid][0] = (lane < ne) ? xr : 0.;
    }

    // J v with the Jacobian columns read from LDS, four independent partial sums so that
    // the loads overlap.
    __device__ __forceinline__ void jac_times_impl(int src, int dst) {
        const double v = V[src][0];
        double a0 = 0., a1 = 0., a2 = 0., a3 = 0.;
        int j = 0;
        for (; j + 4 <= ne; j += 4) {
            const double c0 = Jl[j * 65 + lane], c1 = Jl[(j + 1) * 65 + lane];
            const double c2 = Jl[(j + 2) * 65 + lane], c3 = Jl[(j + 3) * 65 + lane];
            a0 = fma(c0, readlane_d(v, j), a0);
            a1 = fma(c1, readlane_d(v, j + 1), a1);
            a2 = fma(c2, readlane_d(v, j + 2), a2);
            a3 = fma(c3, readlane_d(v, j + 3), a3);
        }
        for (; j < ne; j++) a0 = fma(Jl[j * 65 + lane], readlane_d(v, j), a0);
        V[dst][0] = (a0 + a1) + (a2 + a3);
    }
    // Returns the number of RHS evaluations the Jacobian took.
    __device__ __forceinline__ int jacobian(double h) {
        LSM_TIC(t0);
#ifdef LSM_PHASE_TIMERS
        in_jac = 1;
#endif
        int nrhs = 0;
        if (K.jac_mode == 1) {
            jacobian_analytic_impl();
        } else {
            jacobian_impl(h);
            nrhs = ne;
        }
#ifdef LSM_PHASE_TIMERS
        in_jac = 0;
#endif
        LSM_TOC(1, t0);
        return nrhs;
    }
    __device__ __forceinline__ void factor(double hr0) { LSM_TIC(t0); factor_impl(hr0); LSM_TOC(2, t0); }
    __device__ __forceinline__ void solve(int vid) { LSM_TIC(t0); solve_impl(vid); LSM_TOC(3, t0); }
    __device__ __forceinline__ void jac_times(int src, int dst) { LSM_TIC(t0); jac_times_impl(src, dst); LSM_TOC(4, t0); }
};


// Up to 4 waves per workgroup; registers capped for TEAM_WAVES_PER_SIMD(NMAX) waves per SIMD.
template <int NMAX, bool XK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TEAM_WAVES_PER_SIMD(NMAX))))
team_kernel(DevMech M, KParams K, int64_t n, const double *__restrict__ dt, const double *yin, double *yout,
            int *status, double *q, int *counters, unsigned *queue, const int *order) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    {   // stage the mechanism image (once per workgroup)
        const int4 *src = (const int4 *)M.img.src;
        int4 *dst = (int4 *)lds_raw;
        const int nv = M.img.bytes / 16;
        for (int i = threadIdx.x; i < nv; i += blockDim.x) dst[i] = src[i];
        __syncthreads();
    }
    const int wave = threadIdx.x >> 6;
    double *wave_lds = (double *)(lds_raw + M.img.bytes + (size_t)wave * team_wave_bytes(M.ne));
    TeamCtx<NMAX, XK> c(M, K, lds_raw, wave_lds);
    const int lane = c.lane, ns = M.ns, ne = M.ne;
    for (;;) {
        unsigned cell_u = 0;
        if (lane == 0) cell_u = atomicAdd(queue, 1u);
        cell_u = __shfl(cell_u, 0, 64);
        const int64_t ticket = (int64_t)__builtin_amdgcn_readfirstlane((int)cell_u);
        if (ticket >= n) break;
        // visiting order (e.g. most expensive cells first); entries outside [0, n) are skipped
        const int64_t cell = order ? (int64_t)__builtin_amdgcn_readfirstlane(order[ticket]) : ticket;
        if (cell < 0 || cell >= n) continue;
        {   // opaque per-cell copies: keep lane/size-derived constants from being hoisted
            // out of the cell loop (they would stay live in registers for the whole kernel)
            int l = threadIdx.x & 63, a = M.ne, b = M.ns, r = M.nr, s0 = M.img.simple0;
            asm volatile("" : "+v"(l));
            asm volatile("" : "+s"(a), "+s"(b), "+s"(r), "+s"(s0));
            c.lane = l; c.ne = a; c.ns = b; c.nr = r; c.s0 = s0;
        }
        LSM_TIC(tcell);
        const double yi = lane < ne ? yin[(int64_t)lane * n + cell] : 0.;
        const double T0 = readlane_d(yi, ns);
        double y;
        int st = 0;
        BdfCounters cnt = BdfCounters{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (T0 > K.min_temperature) {                                  // chemistry_DI.H:97
            double y0 = yi;                                            // :101-128
            if (lane < ns && y0 < 0.) y0 = 0.;
            const double sum = wave_sum(lane < ns ? y0 : 0.);
            if (lane < ns) y0 /= sum;
            c.V[Y0_][0] = lane < ne ? y0 : 0.;
            if (K.const_volume) c.set_density(y0, T0, K.rho, cell);     // :173
            st = bdf_integrate(c, K.bk, K.bk.t0 + dt[cell], cnt);       // :138-157
            y = st > 0 ? c.V[Y_][0] : c.V[Y0_][0];
            if (st > 0 && q != nullptr && lane == 0) q[cell] = c.qr;    // :168
        } else {
            y = yi;
        }
        const double sum = wave_sum(lane < ns ? y : 0.);               // :219 normalizeMassFractions
        if (lane < ns) y /= sum;
        if (!K.energy && lane == ns) y = yi;
        if (K.compact) y = c.compact_sources(y, yi, T0, dt[cell]);     // :236-282
        if (lane < ne) yout[(int64_t)lane * n + cell] = y;
        if (lane == 0) {
            if (status) status[cell] = st;
        }
#ifdef LSM_PHASE_TIMERS
        c.tp[5] += clock64() - tcell;
        c.tp[6] += 1;
#endif
        if (counters != nullptr && lane < LSM_NCOUNTERS_DEV) {
            const unsigned vals[12] = {cnt.steps, cnt.fcalls, cnt.fcalls_jac, cnt.jacobians, cnt.factorizations,
                                       cnt.solves, cnt.conv_failures, cnt.err_failures, cnt.max_order,
                                       cnt.rhs_total, cnt.clamps, cnt.newton};
            unsigned v = 0;
#pragma unroll
            for (int i = 0; i < 12; i++)
                if (lane == i) v = vals[i];
            counters[cell * LSM_NCOUNTERS_DEV + lane] = (int)v;
        }
    }
#ifdef LSM_PHASE_TIMERS
    if (lane == 0 && K.phase_cycles)
        for (int i = 0; i < 12; i++) atomicAdd(&K.phase_cycles[i], c.tp[i]);
#endif
}

// ---------------------------------------------------------------------------
// Packed context: CPW = 64 / SEG cells per wavefront, SEG = 16 (ns + 1 <= 16) or 32
// (ns + 1 <= 32).  The wave is split into segments of SEG lanes; segment s integrates
// its own cell, element e of every vector on lane s*SEG + e.  The arithmetic per cell
// is the team context's, operation for operation; what changes is how values move
// between the lanes of one cell:
//   * sums: the team's row scan (row_shr 1, 2, 4, 8; for SEG = 32 row_bcast:15 into the
//     odd row), so a sum over a cell adds in the team's order, then the segment's last
//     lane is broadcast (DPP row_newbcast, and v_permlane16_swap across the two rows of
//     a 32-lane segment);
//   * element j of a cell to all its lanes, j known at compile time (solves, J*v): DPP
//     row_newbcast:j (+ permlane16_swap), no LDS and no scalar unit;
//   * element j at run time (T, FD columns, the pivot-order permutation): ds_bpermute;
//   * the pivot row of an elimination step: written by the pivot lane to a per-cell LDS
//     row and read back by the segment.
// After each factorisation the rows are permuted into pivot order (lane k holds the row
// chosen at step k), so that every broadcast of the triangular solves is a DPP move from a
// compile-time lane.  The BDF driver runs per segment: its scalars are per-lane values,
// uniform within a segment, and segments of one wave branch apart under the exec mask.
// Per-cell LDS: (c, g) pairs and h [64] (slot 63 is the dummy species of fixed-format
// records), R, dR/dT, collider terms and the pivot row [SEG] each, then the Jacobian, ne
// columns of SEG + 1 doubles (lane k's row at offset k).
// ---------------------------------------------------------------------------
template <int SEG>
__device__ __forceinline__ long long seg_bcast_last_i64(long long v) {
    // the last lane of each segment to the whole segment
    int lo = __builtin_amdgcn_update_dpp(0, (int)(v & 0xffffffffll), 0x15F, 0xf, 0xf, false);   // row_newbcast:15
    int hi = __builtin_amdgcn_update_dpp(0, (int)(v >> 32), 0x15F, 0xf, 0xf, false);
    if (SEG == 32) {   // odd rows hold the segment's value: copy them onto the even rows
        lo = (int)__builtin_amdgcn_permlane16_swap((unsigned)lo, (unsigned)lo, false, false)[1];
        hi = (int)__builtin_amdgcn_permlane16_swap((unsigned)hi, (unsigned)hi, false, false)[1];
    }
    return ((long long)hi << 32) | (unsigned)lo;
}
template <int SEG>
__device__ __forceinline__ double seg_bcast_last(double x) {
    return __longlong_as_double(seg_bcast_last_i64<SEG>(__double_as_longlong(x)));
}

// Element J (compile time) of each segment to the whole segment.
template <int SEG, int J>
__device__ __forceinline__ double seg_bcast(double x) {
    static_assert(J >= 0 && J < SEG, "lane outside the segment");
    const long long v = __double_as_longlong(x);
    int lo = __builtin_amdgcn_update_dpp(0, (int)(v & 0xffffffffll), 0x150 + (J & 15), 0xf, 0xf, false);
    int hi = __builtin_amdgcn_update_dpp(0, (int)(v >> 32), 0x150 + (J & 15), 0xf, 0xf, false);
    if (SEG == 32) {   // [0]: even rows copied onto odd rows, [1]: odd rows onto even rows
        lo = (int)__builtin_amdgcn_permlane16_swap((unsigned)lo, (unsigned)lo, false, false)[J < 16 ? 0 : 1];
        hi = (int)__builtin_amdgcn_permlane16_swap((unsigned)hi, (unsigned)hi, false, false)[J < 16 ? 0 : 1];
    }
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Value of lane `src` (a wave lane index, any at run time) on every lane.
__device__ __forceinline__ double lane_shfl_d(double x, int src) {
    const long long v = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_ds_bpermute(src << 2, (int)(v & 0xffffffffll));
    const int hi = __builtin_amdgcn_ds_bpermute(src << 2, (int)(v >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over each segment, in the order of wave_sum restricted to the segment's lanes.
template <int SEG>
__device__ __forceinline__ double seg_sum(double x) {
    x += dpp0_d<0x111, 0xf>(x);
    x += dpp0_d<0x112, 0xf>(x);
    x += dpp0_d<0x114, 0xf>(x);
    x += dpp0_d<0x118, 0xf>(x);
    if (SEG == 32) x += dpp0_d<0x142, 0xa>(x);
    return seg_bcast_last<SEG>(x);
}
template <int SEG>
__device__ __forceinline__ void seg_sum2(double &a, double &b) {
    a += dpp0_d<0x111, 0xf>(a); b += dpp0_d<0x111, 0xf>(b);
    a += dpp0_d<0x112, 0xf>(a); b += dpp0_d<0x112, 0xf>(b);
    a += dpp0_d<0x114, 0xf>(a); b += dpp0_d<0x114, 0xf>(b);
    a += dpp0_d<0x118, 0xf>(a); b += dpp0_d<0x118, 0xf>(b);
    if (SEG == 32) { a += dpp0_d<0x142, 0xa>(a); b += dpp0_d<0x142, 0xa>(b); }
    a = seg_bcast_last<SEG>(a);
    b = seg_bcast_last<SEG>(b);
}
// Maximum over each segment of non-negative doubles.
template <int SEG>
__device__ __forceinline__ double seg_max_nonneg(double x) {
    x = fmax(x, dpp0_d<0x111, 0xf>(x));
    x = fmax(x, dpp0_d<0x112, 0xf>(x));
    x = fmax(x, dpp0_d<0x114, 0xf>(x));
    x = fmax(x, dpp0_d<0x118, 0xf>(x));
    if (SEG == 32) x = fmax(x, dpp0_d<0x142, 0xa>(x));
    return seg_bcast_last<SEG>(x);
}

// Per-cell LDS of the packed context, in doubles: 128 (c, g), 64 h, 4 x SEG (R, dR/dT,
// collider terms, pivot row), the Jacobian ne x (SEG + 1); even, so that every cell's pairs
// stay 16-byte aligned.
__host__ __device__ inline size_t pack_cell_doubles(int seg, int ne) {
    return ((size_t)192 + 4 * (size_t)seg + (size_t)ne * (seg + 1) + 1) / 2 * 2;
}
__host__ __device__ inline size_t pack_wave_bytes(int seg, int ne) {
    return (size_t)(64 / seg) * pack_cell_doubles(seg, ne) * 8;
}

template <int SEG, bool XK>
struct SegCtx {
    static constexpr int JS = SEG + 1;              // Jacobian column stride
    static constexpr unsigned long long SEGMASK = SEG == 64 ? ~0ull : ((1ull << SEG) - 1ull);
    double V[NVEC][1];
    double G[SEG];   // this lane's row of the LU factors; after a factorisation, the row chosen at
                     // elimination step el (multipliers in columns < el, U from el on)
    int ps;          // elimination step at which this lane's row became the pivot (during factor)
    int plv;         // row chosen as pivot at step el
    double dinv;     // 1 / pivot of this lane's row
    double *Jl;      // per-cell LDS: Jacobian [ne][SEG + 1], lane el holds row el
    double2 *scg;    // per-cell LDS: (c_k, g_k)
    double *sR, *sDT, *sH, *sU, *prow;
    LdsMech L;
    const KParams &K;
    int el, sbase, ne, ns, nr;
    int s0;
    double qr;
    double rho0;
#ifdef LSM_PHASE_TIMERS
    __device__ __forceinline__ void add_time(int, unsigned long long) {}
#endif

    __device__ SegCtx(const DevMech &m, const KParams &k, const unsigned char *img, double *cell_lds)
        : L(img, m.img, m.ns_pad), K(k), el(threadIdx.x & (SEG - 1)), sbase(threadIdx.x & 63 & ~(SEG - 1)),
          ne(m.ne), ns(m.ns), nr(m.nr), s0(m.img.simple0), qr(0.), rho0(0.) {
        scg = (double2 *)cell_lds;
        sH = cell_lds + 128;
        sR = sH + 64;
        sDT = sR + SEG;
        sU = sDT + SEG;
        prow = sU + SEG;
        Jl = prow + SEG;
    }

    template <class F>
    __device__ __forceinline__ void each(F f) { f(0); }
    __device__ __forceinline__ double &x(int vid, int e) { return V[vid][e]; }
    __device__ __forceinline__ double ymin(int) const { return el < ns ? 0. : (el == ns ? 200. : NEG_INF); }
    __device__ __forceinline__ double ymax(int) const { return el < ns ? 1. : (el == ns ? 6000. : POS_INF); }
    __device__ __forceinline__ double reduce(double v) { return seg_sum<SEG>(v); }
    __device__ __forceinline__ int count(bool b) {
        return __popcll((__ballot(b) >> sbase) & SEGMASK);
    }
    // element j (run time) of this cell
    __device__ __forceinline__ double elem(double v, int j) const { return lane_shfl_d(v, sbase + j); }

    __device__ __forceinline__ double *racc(unsigned k) const {
        return k == (unsigned)LSM_DUMMY_SP ? sU + el : sR + k;
    }
    __device__ __forceinline__ void reaction(int j, double T, double logT, double uT, double logPRT, double Pk,
                                            double cTot) {
        rxn_rate<XK>(L, scg, j, T, logT, uT, logPRT, Pk, cTot, [&](unsigned k, double v) { atomicAdd(racc(k), v); });
    }
    // Two fixed-format records without collider, reactions j and j + SEG (the team context's
    // simple_pair with the segment's stride).
    __device__ __forceinline__ void simple_pair(int j, double logT, double uT, double logPRT) {
        const int j2 = j + SEG;
        const bool has2 = j2 < nr;
        const int jb = has2 ? j2 : j;
        const uint4 w1 = L.W[j], w2 = L.W[jb];
        const double2 ab1 = ((const double2 *)L.P)[2 * j], ed1 = ((const double2 *)L.P)[2 * j + 1];
        const double2 ab2 = ((const double2 *)L.P)[2 * jb], ed2 = ((const double2 *)L.P)[2 * jb + 1];
        const unsigned r10 = w1.y & 0xffu, r11 = (w1.y >> 8) & 0xffu, r12 = (w1.y >> 16) & 0xffu;
        const unsigned p10 = w1.z & 0xffu, p11 = (w1.z >> 8) & 0xffu, p12 = (w1.z >> 16) & 0xffu;
        const unsigned r20 = w2.y & 0xffu, r21 = (w2.y >> 8) & 0xffu, r22 = (w2.y >> 16) & 0xffu;
        const unsigned p20 = w2.z & 0xffu, p21 = (w2.z >> 8) & 0xffu, p22 = (w2.z >> 16) & 0xffu;
        const double2 a10 = scg[r10], a11 = scg[r11], a12 = scg[r12], b10 = scg[p10], b11 = scg[p11], b12 = scg[p12];
        const double2 a20 = scg[r20], a21 = scg[r21], a22 = scg[r22], b20 = scg[p20], b21 = scg[p21], b22 = scg[p22];
        double kf1 = exp(ab1.x + ab1.y * logT - ed1.x * uT);
        double kf2 = exp(ab2.x + ab2.y * logT - ed2.x * uT);
        kf1 = (w1.x & 2u) ? -kf1 : kf1;
        kf2 = (w2.x & 2u) ? -kf2 : kf2;
        const double dg1 = (b10.y + b11.y + b12.y) - (a10.y + a11.y + a12.y);
        const double dg2 = (b20.y + b21.y + b22.y) - (a20.y + a21.y + a22.y);
        const double uk1 = exp(dg1 - logPRT * ed1.y), uk2 = exp(dg2 - logPRT * ed2.y);
        double net1 = (a10.x * a11.x) * a12.x, net2 = (a20.x * a21.x) * a22.x;
        const double rb1 = ((b10.x * b11.x) * b12.x) * uk1, rb2 = ((b20.x * b21.x) * b22.x) * uk2;
        net1 = (w1.x & 1u) ? net1 - rb1 : net1;
        net2 = (w2.x & 1u) ? net2 - rb2 : net2;
        const double r1 = net1 * kf1;
        const double r2 = has2 ? net2 * kf2 : 0.;
        atomicAdd(racc(r10), -r1); atomicAdd(racc(r11), -r1); atomicAdd(racc(r12), -r1);
        atomicAdd(racc(p10), r1);  atomicAdd(racc(p11), r1);  atomicAdd(racc(p12), r1);
        atomicAdd(racc(r20), -r2); atomicAdd(racc(r21), -r2); atomicAdd(racc(r22), -r2);
        atomicAdd(racc(p20), r2);  atomicAdd(racc(p21), r2);  atomicAdd(racc(p22), r2);
    }

    // BatchReactorHomogeneousConstantPressure::Equations (:175-247) / ConstantVolume::Equations
    // (BatchReactorHomogeneousConstantVolume.H:150-199), one cell per segment.
    __device__ __forceinline__ void rhs(int src, int dst) {
        const double T = elem(V[src][0], ns);
        const int k = el;
        const bool isp = k < ns;
        const double Yk = isp ? V[src][0] : 0.;
        const int P = L.SP;
        const double mwk = isp ? L.sp[15 * P + k] : 1.;
        const double MW = 1. / seg_sum<SEG>(Yk / mwk);
        const double xk = (MW * Yk) / mwk;
        const double cTotP = K.const_volume ? rho0 / MW : K.pressure / RGAS / T;
        const double ck = cTotP * xk;
        const double rho = K.const_volume ? rho0 : cTotP * MW;
        const double T2 = T * T, T3 = T2 * T, T4 = T3 * T, uT = 1. / T, logT = log(T);
        double cp = 0., h = 0.;
        if (isp) {
            const double *a = L.sp + ((T > L.sp[14 * P + k]) ? 0 : 7 * P) + k;
            const double a0 = a[0], a1 = a[P], a2 = a[2 * P], a3 = a[3 * P], a4 = a[4 * P];
            cp = a0 + T * a1 + T2 * a2 + T3 * a3 + T4 * a4;
            h = a0 + T * (0.50 * a1) + T2 * ((1. / 3.) * a2) + T3 * (0.25 * a3) + T4 * (0.20 * a4) + uT * a[5 * P];
            const double s = a0 * logT + T * a1 + T2 * (0.50 * a2) + T3 * ((1. / 3.) * a3) + T4 * (0.25 * a4) +
                             a[6 * P];
            scg[k] = make_double2(ck, h - s);
        }
        if (k == SEG - 1) scg[LSM_DUMMY_SP] = make_double2(1., 0.);
        sR[k] = 0.;
        double cTot = ck, cpx = isp ? cp * xk : 0.;
        seg_sum2<SEG>(cTot, cpx);
        wave_sync();
        const double logPRT = log(101325. / RGAS / T);
        const double Pk = K.const_volume ? cTotP * RGAS * T : K.pressure;
        for (int j = el; j < s0; j += SEG) reaction(j, T, logT, uT, logPRT, Pk, cTot);
        for (int j = s0 + el; j < nr; j += 2 * SEG) simple_pair(j, logT, uT, logPRT);
        wave_sync();
        const double Rk = isp ? sR[k] : 0.;
        double dy = isp ? (mwk * Rk) / rho : 0.;
        if (K.energy && K.const_volume) {
            double hs = isp ? Rk * h : 0., rs = Rk;
            seg_sum2<SEG>(hs, rs);
            qr = -hs * RGAS * T;
            const double CvMixMass = (cpx * RGAS - RGAS) / MW;
            if (k == ns) dy = (qr + RGAS * T * rs) / (rho * CvMixMass);
        } else if (K.energy) {
            const double hs = seg_sum<SEG>(isp ? Rk * h : 0.);
            qr = -hs * RGAS * T;
            const double CpMixMass = (cpx * RGAS) / MW;
            if (k == ns) dy = qr / (rho * CpMixMass);
        }
        wave_sync();
        V[dst][0] = dy;
    }

    __device__ __forceinline__ void set_density(double y0, double T0, const double *rho_cells, int64_t cell) {
        const double mwk = el < ns ? L.sp[15 * L.SP + el] : 1.;
        const double MW0 = 1. / seg_sum<SEG>(el < ns ? y0 / mwk : 0.);
        rho0 = rho_cells ? rho_cells[cell] : K.pressure * MW0 / (RGAS * T0);
    }

    __device__ __forceinline__ double compact_sources(double y, double yin, double Tin, double dtc) {
        const bool isp = el < ns;
        const double Tf = elem(y, ns);
        const int P = L.SP;
        const double mwk = isp ? L.sp[15 * P + el] : 1.;
        double cp = 0.;
        if (isp) {
            const double *a = L.sp + ((Tf > L.sp[14 * P + el]) ? 0 : 7 * P) + el;
            cp = a[0] + Tf * (a[P] + Tf * (a[2 * P] + Tf * (a[3 * P] + Tf * a[4 * P])));
        }
        double u = isp ? y / mwk : 0., cy = isp ? cp * (y / mwk) : 0.;
        seg_sum2<SEG>(u, cy);
        const double rho = K.pressure * (1. / u) / (RGAS * Tf);
        const bool ok = dtc > 1.e-14;
        if (isp) return ok ? rho * (y - yin) / dtc : 0.;
        return (ok && K.energy) ? rho * (cy * RGAS) * (Tf - Tin) / dtc : 0.;
    }

    struct JacAcc {
        double *sR, *sDT, *Dm, *sU;
        __device__ __forceinline__ void R(unsigned k, double r, double drdT) const {
            if (k != (unsigned)LSM_DUMMY_SP) {
                atomicAdd(&sR[k], r);
                atomicAdd(&sDT[k], drdT);
            }
        }
        __device__ __forceinline__ void D(unsigned i, unsigned k, double v) const { atomicAdd(&Dm[i * JS + k], v); }
        __device__ __forceinline__ void U(unsigned k, double v) const { atomicAdd(&sU[k], v); }
    };

    // The team context's analytic Jacobian (same formulas, same order), one cell per segment.
    __device__ __forceinline__ void jacobian_analytic_impl() {
        const double T = elem(V[VB_][0], ns);
        const int k = el;
        const bool isp = k < ns;
        const double Yk = isp ? V[VB_][0] : 0.;
        const int P = L.SP;
        const double mwk = isp ? L.sp[15 * P + k] : 1.;
        const double MW = 1. / seg_sum<SEG>(Yk / mwk);
        const double xk = (MW * Yk) / mwk;
        const double cTotP = K.pressure / RGAS / T;
        const double ck = cTotP * xk;
        const double rho = cTotP * MW;
        const double T2 = T * T, T3 = T2 * T, T4 = T3 * T, uT = 1. / T, logT = log(T);
        double cp = 0., h = 0., dcp = 0.;
        if (isp) {
            const double *a = L.sp + ((T > L.sp[14 * P + k]) ? 0 : 7 * P) + k;
            const double a0 = a[0], a1 = a[P], a2 = a[2 * P], a3 = a[3 * P], a4 = a[4 * P];
            cp = a0 + T * a1 + T2 * a2 + T3 * a3 + T4 * a4;
            dcp = a1 + T * (2. * a2) + T2 * (3. * a3) + T3 * (4. * a4);
            h = a0 + T * (0.50 * a1) + T2 * ((1. / 3.) * a2) + T3 * (0.25 * a3) + T4 * (0.20 * a4) + uT * a[5 * P];
            const double s = a0 * logT + T * a1 + T2 * (0.50 * a2) + T3 * ((1. / 3.) * a3) + T4 * (0.25 * a4) +
                             a[6 * P];
            scg[k] = make_double2(ck, h - s);
            sH[k] = h;
        }
        if (k == SEG - 1) { scg[LSM_DUMMY_SP] = make_double2(1., 0.); sH[LSM_DUMMY_SP] = 0.; }
        sR[k] = 0.;
        sDT[k] = 0.;
        sU[k] = 0.;
        for (int i = 0; i < ne; i++) Jl[i * JS + k] = 0.;
        double cTot = ck, cpx = isp ? cp * xk : 0.;
        seg_sum2<SEG>(cTot, cpx);
        wave_sync();
        const double logPRT = log(101325. / RGAS / T);
        const JacAcc acc{sR, sDT, Jl, sU};
        for (int j = el; j < nr; j += SEG)
            rxn_jac<XK>(L, scg, sH, ns, LSM_DUMMY_SP, j, T, logT, uT, logPRT, K.pressure, cTot, acc);
        wave_sync();
        const double Rk = isp ? sR[k] : 0.;
        const double dRdT = isp ? sDT[k] : 0.;
        const double Uk = isp ? sU[k] : 0.;
        double S = 0., A = 0.;
        if (isp) {
            for (int i = 0; i < ns; i++) S += (Jl[i * JS + k] + Uk) * scg[i].x;
            for (int i = 0; i < ns; i++) A += sH[i] * (Jl[k * JS + i] + sU[i]);
        }
        wave_sync();
        const double fk = isp ? (mwk * Rk) / rho : 0.;
        if (isp) {
            const double g = fk - mwk * S / rho;
            for (int j = 0; j < ns; j++) {
                const double mwj = L.sp[15 * P + j];
                Jl[j * JS + k] = (mwk / mwj) * (Jl[j * JS + k] + Uk) + (MW / mwj) * g;
            }
            Jl[ns * JS + k] = (mwk / rho) * (dRdT - S * uT) + fk * uT;
        } else if (k > ns) {
            for (int j = 0; j < ne; j++) Jl[j * JS + k] = 0.;
        }
        const double cpm = (cpx * RGAS) / MW;
        const double Phi = rho * cpm;
        double hs = isp ? h * Rk : 0., hS = isp ? h * S : 0.;
        seg_sum2<SEG>(hs, hS);
        const double QR = -hs * RGAS * T;
        const double fT = K.energy ? QR / Phi : 0.;
        double dQ = isp ? -(cp * RGAS * Rk + h * RGAS * T * (dRdT - S * uT)) : 0.;
        double dP = isp ? Yk * dcp * RGAS / mwk : 0.;
        seg_sum2<SEG>(dQ, dP);
        if (K.energy) {
            if (isp) {
                const double dQdY = -(RGAS * T / mwk) * (rho * A - MW * hS);
                const double dPdY = (rho / mwk) * (cp * RGAS - MW * cpm);
                Jl[k * JS + ns] = (dQdY - fT * dPdY) / Phi;
            }
            if (k == ns) Jl[ns * JS + ns] = (dQ - fT * (-Phi * uT + rho * dP)) / Phi;
        } else {
            if (isp) Jl[k * JS + ns] = 0.;
            if (k == ns) Jl[ns * JS + ns] = 0.;
        }
        wave_sync();
    }

    // KernelDense::NumericalJacobian, max-constraint branch (KernelDense.hpp:160-183)
    __device__ __forceinline__ void jacobian_impl(double h) {
        const double ZERO_DER = sqrt(TINY_F), ETA2 = sqrt(EPS_D), BETA = 1.e+3 * EPS_D;
        double fw = V[F_][0] * V[OOE_][0];
        double hf = BETA * fabs(h) * sqrt(seg_sum<SEG>(fw * fw) / (double)ne) * (double)ne;
        if (hf < 1.e-10) hf = 1.;
        for (int j = 0; j < ne; j++) {
            const double yh = elem(V[VB_][0], j);
            const double ej = elem(V[OOE_][0], j);
            const double a = fabs(yh), b = fabs(1. / ej);
            double hJ = ETA2 * (a > b ? a : b);
            hJ = fmax(hJ, hf / ej);
            hJ = fmax(hJ, ZERO_DER);
            const double ymx = j < ns ? 1. : 6000.;
            if (yh + hJ > ymx) hJ = -hJ;
            if (el == j) V[VB_][0] = yh + hJ;
            rhs(VB_, AUX_);
            if (el == j) V[VB_][0] = yh;
            const double hInv = 1. / hJ;
            Jl[j * JS + el] = hInv * (V[AUX_][0] - V[F_][0]);
        }
    }

    // Element of the largest |g| among the eligible elements of this cell (the lowest on a tie;
    // the first eligible one when every eligible |g| is 0).
    __device__ __forceinline__ int pivot_el(double g, bool elig) {
        const double a = elig ? fabs(g) : 0.;
        const double m = seg_max_nonneg<SEG>(a);
        const unsigned long long hit = (__ballot(elig && a == m) >> sbase) & SEGMASK;
        return __ffsll(hit) - 1;
    }

    // G = I - hr0 J and row-pivoted LU (KernelDense.hpp:196-221), the team context's
    // elimination with the pivot row staged in LDS; the rows are then permuted into pivot order.
    __device__ __forceinline__ void factor_impl(double hr0) {
        int el = this->el, ne = this->ne;
        asm volatile("" : "+v"(el));
        asm volatile("" : "+s"(ne));
#pragma unroll
        for (int j = 0; j < SEG; j++) {
            const double jij = (j < ne) ? Jl[j * JS + el] : 0.;
            G[j] = jij * -hr0 + ((el == j) ? 1. : 0.);
        }
        ps = (el < ne) ? -1 : 1 << 20;
        dinv = 1.;
        int p = pivot_el(G[0], ps < 0);
#pragma unroll
        for (int k = 0; k < SEG; k++) {
            if (k < ne) {
                if (el == p) {
#pragma unroll
                    for (int j = k; j < SEG; j++) prow[j] = G[j];
                }
                wave_sync();
                const double piv = prow[k];
                const double inv = 1. / piv;
                if (el == p) { ps = k; dinv = inv; }
                double l = 0.;
                if (ps < 0 && piv != 0.) {
                    l = G[k] * inv;
                    G[k] = l;
                }
                int pn = 0;
                if (k + 1 < SEG) {
                    G[k + 1] = fma(-l, prow[k + 1], G[k + 1]);
                    pn = pivot_el(G[k + 1], ps < 0);
                }
#pragma unroll
                for (int j = k + 2; j < SEG; j++) G[j] = fma(-l, prow[j], G[j]);
                wave_sync();   // the pivot row is read before the next step overwrites it
                p = pn;
            }
        }
        // pivot order: element k learns which row was chosen at step k, then takes that row
        // (a column with no eligible finite entry leaves ps at -1: its slot is skipped, the masks
        // keep every index inside the segment, and the NaNs fail the step as in the team context)
        int *pl = (int *)prow;
        if (el < ne && ps >= 0) pl[ps] = el;
        wave_sync();
        plv = (el < ne) ? (pl[el] & (SEG - 1)) : el;
        wave_sync();
        const int src = sbase + plv;
#pragma unroll
        for (int j = 0; j < SEG; j++) G[j] = lane_shfl_d(G[j], src);
        dinv = lane_shfl_d(dinv, src);
    }

    // Forward / backward substitution with the rows in pivot order: the broadcasts of the
    // dependent chain come from compile-time lanes (DPP).  Same operations and order as the
    // team context's solve, in groups of 8 steps.
    template <int K0>
    __device__ __forceinline__ void fwd_group(double &b, int ne, int el_l) const {
        if (K0 < ne) {
            double lk[8];
#pragma unroll
            for (int u = 0; u < 8; u++) lk[u] = G[K0 + u] * ((el_l > K0 + u) ? 1. : 0.);
            b = fma(-lk[0], seg_bcast<SEG, K0 + 0>(b), b);
            b = fma(-lk[1], seg_bcast<SEG, K0 + 1>(b), b);
            b = fma(-lk[2], seg_bcast<SEG, K0 + 2>(b), b);
            b = fma(-lk[3], seg_bcast<SEG, K0 + 3>(b), b);
            b = fma(-lk[4], seg_bcast<SEG, K0 + 4>(b), b);
            b = fma(-lk[5], seg_bcast<SEG, K0 + 5>(b), b);
            b = fma(-lk[6], seg_bcast<SEG, K0 + 6>(b), b);
            b = fma(-lk[7], seg_bcast<SEG, K0 + 7>(b), b);
        }
    }
    template <int K0>
    __device__ __forceinline__ void bwd_group(double &b, double &xr, double dinv_l, int ne, int el_l) const {
        if (K0 < ne) {
            double uk[8], dk[8];
#pragma unroll
            for (int u = 0; u < 8; u++) uk[u] = G[K0 + u] * ((el_l < K0 + u) ? 1. : 0.);
            dk[7] = seg_bcast<SEG, K0 + 7>(dinv_l);
            dk[6] = seg_bcast<SEG, K0 + 6>(dinv_l);
            dk[5] = seg_bcast<SEG, K0 + 5>(dinv_l);
            dk[4] = seg_bcast<SEG, K0 + 4>(dinv_l);
            dk[3] = seg_bcast<SEG, K0 + 3>(dinv_l);
            dk[2] = seg_bcast<SEG, K0 + 2>(dinv_l);
            dk[1] = seg_bcast<SEG, K0 + 1>(dinv_l);
            dk[0] = seg_bcast<SEG, K0 + 0>(dinv_l);
            double xk;
            xk = seg_bcast<SEG, K0 + 7>(b) * dk[7]; b = fma(-uk[7], xk, b); if (el_l == K0 + 7) xr = xk;
            xk = seg_bcast<SEG, K0 + 6>(b) * dk[6]; b = fma(-uk[6], xk, b); if (el_l == K0 + 6) xr = xk;
            xk = seg_bcast<SEG, K0 + 5>(b) * dk[5]; b = fma(-uk[5], xk, b); if (el_l == K0 + 5) xr = xk;
            xk = seg_bcast<SEG, K0 + 4>(b) * dk[4]; b = fma(-uk[4], xk, b); if (el_l == K0 + 4) xr = xk;
            xk = seg_bcast<SEG, K0 + 3>(b) * dk[3]; b = fma(-uk[3], xk, b); if (el_l == K0 + 3) xr = xk;
            xk = seg_bcast<SEG, K0 + 2>(b) * dk[2]; b = fma(-uk[2], xk, b); if (el_l == K0 + 2) xr = xk;
            xk = seg_bcast<SEG, K0 + 1>(b) * dk[1]; b = fma(-uk[1], xk, b); if (el_l == K0 + 1) xr = xk;
            xk = seg_bcast<SEG, K0 + 0>(b) * dk[0]; b = fma(-uk[0], xk, b); if (el_l == K0 + 0) xr = xk;
        }
    }
    template <int... I>
    __device__ __forceinline__ void fwd_all(double &b, int ne, int el_l, std::integer_sequence<int, I...>) const {
        (fwd_group<8 * I>(b, ne, el_l), ...);                       // groups 0, 8, ...
    }
    template <int... I>
    __device__ __forceinline__ void bwd_all(double &b, double &xr, double dinv_l, int ne, int el_l,
                                            std::integer_sequence<int, I...>) const {
        (bwd_group<SEG - 8 - 8 * I>(b, xr, dinv_l, ne, el_l), ...);  // groups SEG - 8, ..., 0
    }
    __device__ __forceinline__ void solve_impl(int vid) {
        int ne = this->ne, el_l = el, src = sbase + plv;
        double dinv_l = dinv;
        asm volatile("" : "+s"(ne));
        asm volatile("" : "+v"(el_l), "+v"(src), "+v"(dinv_l));
        double b = lane_shfl_d(V[vid][0], src);   // b of the row chosen at step el
        fwd_all(b, ne, el_l, std::make_integer_sequence<int, SEG / 8>{});
        double xr = 0.;
        bwd_all(b, xr, dinv_l, ne, el_l, std::make_integer_sequence<int, SEG / 8>{});
        V[vid][0] = (el_l < ne) ? xr : 0.;
    }

    // J v, the team context's four partial sums (element j's column to a[j % 4] for j below the
    // largest multiple of 4 not above ne, the rest to a[0]), with v_j broadcast by DPP.
    template <int J>
    __device__ __forceinline__ void jv_term(double v, int nq, double &a0, double &a1, double &a2, double &a3) const {
        if (J < ne) {
            const double c = Jl[J * JS + el];
            const double vj = seg_bcast<SEG, J>(v);
            if (J < nq) {
                if ((J & 3) == 0) a0 = fma(c, vj, a0);
                else if ((J & 3) == 1) a1 = fma(c, vj, a1);
                else if ((J & 3) == 2) a2 = fma(c, vj, a2);
                else a3 = fma(c, vj, a3);
            } else {
                a0 = fma(c, vj, a0);
            }
        }
    }
    template <int... J>
    __device__ __forceinline__ void jv_all(double v, int nq, double &a0, double &a1, double &a2, double &a3,
                                           std::integer_sequence<int, J...>) const {
        (jv_term<J>(v, nq, a0, a1, a2, a3), ...);
    }
    __device__ __forceinline__ void jac_times_impl(int src, int dst) {
        const double v = V[src][0];
        double a0 = 0., a1 = 0., a2 = 0., a3 = 0.;
        jv_all(v, ne & ~3, a0, a1, a2, a3, std::make_integer_sequence<int, SEG>{});
        V[dst][0] = (a0 + a1) + (a2 + a3);
    }

    __device__ __forceinline__ int jacobian(double h) {
        if (K.jac_mode == 1) { jacobian_analytic_impl(); return 0; }
        jacobian_impl(h);
        return ne;
    }
    __device__ __forceinline__ void factor(double hr0) { factor_impl(hr0); }
    __device__ __forceinline__ void solve(int vid) { solve_impl(vid); }
    __device__ __forceinline__ void jac_times(int src, int dst) { jac_times_impl(src, dst); }
};

// Persistent packed kernel: each wave claims CPW consecutive tickets of the queue, segment s
// integrates the cell of ticket base + s.  The per-cell harness is the team kernel's
// (chemistry_DI.H:97-282).
#define PACK_WAVES_PER_SIMD 1
template <int SEG, bool XK>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PACK_WAVES_PER_SIMD)))
pack_kernel(DevMech M, KParams K, int64_t n, const double *__restrict__ dt, const double *yin, double *yout,
            int *status, double *q, int *counters, unsigned *queue, const int *order) {
    constexpr int CPW = 64 / SEG;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    {
        const int4 *src = (const int4 *)M.img.src;
        int4 *dst = (int4 *)lds_raw;
        const int nv = M.img.bytes / 16;
        for (int i = threadIdx.x; i < nv; i += blockDim.x) dst[i] = src[i];
        __syncthreads();
    }
    const int wave = threadIdx.x >> 6;
    const int sg = (threadIdx.x & 63) / SEG;
    double *cell_lds = (double *)(lds_raw + M.img.bytes + (size_t)wave * pack_wave_bytes(SEG, M.ne)) +
                       (size_t)sg * pack_cell_doubles(SEG, M.ne);
    SegCtx<SEG, XK> c(M, K, lds_raw, cell_lds);
    const int ns = M.ns, ne = M.ne;
    for (;;) {
        unsigned base_u = 0;
        if ((threadIdx.x & 63) == 0) base_u = atomicAdd(queue, (unsigned)CPW);
        base_u = __shfl(base_u, 0, 64);
        const int64_t base = (int64_t)__builtin_amdgcn_readfirstlane((int)base_u);
        if (base >= n) break;
        const int64_t ticket = base + sg;
        int64_t cell = -1;
        if (ticket < n) cell = order ? (int64_t)order[ticket] : ticket;
        if (cell >= 0 && cell < n) {
            {
                int l = threadIdx.x & (SEG - 1), a = M.ne, b = M.ns, r = M.nr, s0 = M.img.simple0;
                asm volatile("" : "+v"(l));
                asm volatile("" : "+s"(a), "+s"(b), "+s"(r), "+s"(s0));
                c.el = l; c.ne = a; c.ns = b; c.nr = r; c.s0 = s0;
            }
            const int el = c.el;
            const double yi = el < ne ? yin[(int64_t)el * n + cell] : 0.;
            const double T0 = c.elem(yi, ns);
            double y;
            int st = 0;
            BdfCounters cnt = BdfCounters{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            if (T0 > K.min_temperature) {                                  // chemistry_DI.H:97
                double y0 = yi;                                            // :101-128
                if (el < ns && y0 < 0.) y0 = 0.;
                const double sum = seg_sum<SEG>(el < ns ? y0 : 0.);
                if (el < ns) y0 /= sum;
                c.V[Y0_][0] = el < ne ? y0 : 0.;
                if (K.const_volume) c.set_density(y0, T0, K.rho, cell);     // :173
                st = bdf_integrate(c, K.bk, K.bk.t0 + dt[cell], cnt);       // :138-157
                y = st > 0 ? c.V[Y_][0] : c.V[Y0_][0];
                if (st > 0 && q != nullptr && el == 0) q[cell] = c.qr;      // :168
            } else {
                y = yi;
            }
            const double sum = seg_sum<SEG>(el < ns ? y : 0.);             // :219
            if (el < ns) y /= sum;
            if (!K.energy && el == ns) y = yi;
            if (K.compact) y = c.compact_sources(y, yi, T0, dt[cell]);     // :236-282
            if (el < ne) yout[(int64_t)el * n + cell] = y;
            if (el == 0 && status) status[cell] = st;
            if (counters != nullptr && el < LSM_NCOUNTERS_DEV) {
                const unsigned vals[12] = {cnt.steps, cnt.fcalls, cnt.fcalls_jac, cnt.jacobians, cnt.factorizations,
                                           cnt.solves, cnt.conv_failures, cnt.err_failures, cnt.max_order,
                                           cnt.rhs_total, cnt.clamps, cnt.newton};
                unsigned v = 0;
#pragma unroll
                for (int i = 0; i < 12; i++)
                    if (el == i) v = vals[i];
                counters[cell * LSM_NCOUNTERS_DEV + el] = (int)v;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Wide context: one cell per 256-thread workgroup (ne <= 256).  Thread e owns
// element e of every vector (registers); sums are wave DPP reductions combined
// across the four waves through LDS (one barrier, fixed order, identical result
// in every thread).  The Newton matrix G = I - h l0 J is factorised in FP32 in
// LDS (row-major, physical row swaps); the Jacobian stays FP64 in an HBM/L2
// workspace, which also serves the FP64 residuals of iterative refinement and
// J*v.  Formation rates are accumulated per wave and summed in wave order, so
// the RHS is bitwise reproducible.
// ---------------------------------------------------------------------------
template <bool XK>
struct WideCtx {
    double V[NVEC][1];
    const KParams &K;
    LdsMech L;
    int tid, wave, lane, ne, ns, nr, lda;
    unsigned dummy;
    double qr;
    double rho0;              // constant-volume reactor: density of the current cell
    double hr0;               // h l0 of the current factorisation (for the residuals)
    int rp;                   // reduction buffer parity
    double2 *scg;             // LDS [256]  (c_k, g_k)
    double *sR, *sDT;         // LDS [4][256] per-wave accumulators
    double *sH;               // LDS [256]  h_k = H_k/RT
    double *red;              // LDS [2][8] reduction / broadcast slots, then the queue slot
    double *bs;               // LDS [256]  solve and mat-vec exchange
    double *dinv;             // LDS [256]  1 / pivot, by row
    int *perm;                // LDS [256]  row chosen as pivot at each elimination step
    unsigned long long *pkey; // LDS [2]    pivot keys (parity of the elimination step)
    float *fsc;               // LDS [1536] factorisation scratch: [0,256) the FP32 solve vector,
                              // [256,384) per-lane targets of the MFMA phases' masked-off stores,
                              // [512,768) the pivot block's inverse, [800,944) pivot-search slots
    float *A;                 // LDS [ne + 1][lda] LU factors (spare last row); FP64 blocks of dR/dc in an analytic Jacobian
    double *J;                // HBM [ne][ne] column-major, J[j*ne + i] = d f_i / d y_j
    double *G64;              // HBM [ne][ne] column-major: FP64 LU of G when refinement does not converge
    int *p64;                 // LDS [256]  row interchanges of the FP64 LU
    bool fp64 = false;        // the current factorisation is the FP64 one
#ifdef LSM_PHASE_TIMERS
    unsigned long long tp[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    __device__ __forceinline__ void add_time(int i, unsigned long long d) { tp[i] += d; }
    int in_jac = 0;
#endif

    __device__ WideCtx(const DevMech &m, const KParams &k, unsigned char *lds, double *jac)
        : K(k), L((const unsigned char *)m.img.src, m.img, m.ns_pad), tid(threadIdx.x), wave(threadIdx.x >> 6),
          lane(threadIdx.x & 63), ne(m.ne), ns(m.ns), nr(m.nr), lda(m.ne | 1), dummy((unsigned)m.img.dummy), qr(0.),
          rho0(0.), hr0(0.), rp(0), J(jac), G64(jac + (size_t)m.ne * m.ne) {
        double *d = (double *)lds;
        scg = (double2 *)d;
        sR = d + 512;
        sDT = d + 1536;
        sH = d + 2560;
        red = d + 2816;             // [2][8] reductions, [16] queue slot
        bs = d + 2840;
        dinv = d + 3096;
        pkey = (unsigned long long *)(d + 3352);
        perm = (int *)(d + 3354);
        fsc = (float *)(d + 3482);  // [1536]
        p64 = (int *)(d + 4250);
        A = (float *)(d + 4378);
    }
    __host__ __device__ static size_t lds_bytes(int ne) { return 8 * 4378 + 4 * (size_t)(ne + 1) * (size_t)(ne | 1); }

    template <class F>
    __device__ __forceinline__ void each(F f) { f(0); }
    __device__ __forceinline__ double &x(int vid, int e) { return V[vid][e]; }
    __device__ __forceinline__ double ymin(int) const {
        return tid < ns ? 0. : (tid == ns ? 200. : NEG_INF);
    }
    __device__ __forceinline__ double ymax(int) const {
        return tid < ns ? 1. : (tid == ns ? 6000. : POS_INF);
    }
    // Sum over the workgroup; slot 4 of the same buffer carries a broadcast value
    // (written by thread `src`, -1 for none).
    __device__ __forceinline__ double wg_sum(double v, double *bc = nullptr, double bv = 0., int src = -1) {
        v = wave_sum(v);
        double *b = red + 8 * rp;
        rp ^= 1;
        if (lane == 0) b[wave] = v;
        if (tid == src) b[4] = bv;
        __syncthreads();
        if (bc) *bc = b[4];
        return (b[0] + b[1]) + (b[2] + b[3]);
    }
    __device__ __forceinline__ void wg_sum2(double &a, double &c) {
        wave_sum2(a, c);
        double *b = red + 8 * rp;
        rp ^= 1;
        if (lane == 0) { b[wave] = a; b[4 + wave] = c; }
        __syncthreads();
        a = (b[0] + b[1]) + (b[2] + b[3]);
        c = (b[4] + b[5]) + (b[6] + b[7]);
    }
    __device__ __forceinline__ double reduce(double v) { return wg_sum(v); }
    __device__ __forceinline__ int count(bool b) { return (int)wg_sum(b ? 1. : 0.); }

    __device__ __forceinline__ void rhs(int src, int dst) {
        LSM_TIC(t0);
        rhs_impl(src, dst);
#ifdef LSM_PHASE_TIMERS
        if (!in_jac) LSM_TOC(0, t0);
#endif
    }

    // NASA-7 of species k at T (thread k): cp/R, h/RT, s/R and d(cp/R)/dT.
    __device__ __forceinline__ void thermo(int k, double T, double &cp, double &h, double &s, double &dcp) const {
        const int P = L.SP;
        const double T2 = T * T, T3 = T2 * T, T4 = T3 * T, uT = 1. / T, logT = log(T);
        const double *a = L.sp + ((T > L.sp[14 * P + k]) ? 0 : 7 * P) + k;
        const double a0 = a[0], a1 = a[P], a2 = a[2 * P], a3 = a[3 * P], a4 = a[4 * P];
        cp = a0 + T * a1 + T2 * a2 + T3 * a3 + T4 * a4;
        dcp = a1 + T * (2. * a2) + T2 * (3. * a3) + T3 * (4. * a4);
        h = a0 + T * (0.50 * a1) + T2 * ((1. / 3.) * a2) + T3 * (0.25 * a3) + T4 * (0.20 * a4) + uT * a[5 * P];
        s = a0 * logT + T * a1 + T2 * (0.50 * a2) + T3 * ((1. / 3.) * a3) + T4 * (0.25 * a4) + a[6 * P];
    }

    // BatchReactorHomogeneousConstantPressure::Equations (:175-247), or with K.const_volume
    // BatchReactorHomogeneousConstantVolume::Equations (BatchReactorHomogeneousConstantVolume.H:150-199)
    __device__ __forceinline__ void rhs_impl(int src, int dst) {
        const int k = tid;
        const bool isp = k < ns;
        const double yv = V[src][0];
        const double mwk = isp ? L.sp[15 * L.SP + k] : 1.;
        double T;
        const double MW = 1. / wg_sum(isp ? yv / mwk : 0., &T, yv, ns);
        const double xk = isp ? (MW * yv) / mwk : 0.;
        const double cTotP = K.const_volume ? rho0 / MW : K.pressure / RGAS / T;
        const double ck = cTotP * xk;
        const double rho = K.const_volume ? rho0 : cTotP * MW;
        double cp = 0., h = 0., s = 0., dcp = 0.;
        if (isp) {
            thermo(k, T, cp, h, s, dcp);
            scg[k] = make_double2(ck, h - s);
        }
        if (k == (int)dummy) scg[k] = make_double2(1., 0.);
#pragma unroll
        for (int w = 0; w < 4; w++) sR[w * 256 + k] = 0.;
        double cTot = ck, cpx = isp ? cp * xk : 0.;
        wg_sum2(cTot, cpx);                        // also publishes scg and the cleared sR
        const double logT = log(T), uT = 1. / T, logPRT = log(101325. / RGAS / T);
        const double Pk = K.const_volume ? cTotP * RGAS * T : K.pressure;   // the kinetics' pressure
        double *sRw = sR + 256 * wave;
        for (int j = tid; j < nr; j += WIDE_THREADS)
            rxn_rate<XK>(L, scg, j, T, logT, uT, logPRT, Pk, cTot, [&](unsigned kk, double v) { atomicAdd(&sRw[kk], v); });
        __syncthreads();
        const double Rk = isp ? ((sR[k] + sR[256 + k]) + (sR[512 + k] + sR[768 + k])) : 0.;
        double dy = isp ? (mwk * Rk) / rho : 0.;
        if (K.energy && K.const_volume) {
            // dT/dt = (QR + R T sum_k R_k) / (rho0 cv), cv = (CpMixMolar - R) / MW
            double hs = isp ? Rk * h : 0., rs = Rk;
            wg_sum2(hs, rs);
            qr = -hs * RGAS * T;
            const double CvMixMass = (cpx * RGAS - RGAS) / MW;
            if (k == ns) dy = (qr + RGAS * T * rs) / (rho * CvMixMass);
        } else if (K.energy) {
            const double hs = wg_sum(isp ? Rk * h : 0.);
            qr = -hs * RGAS * T;
            const double CpMixMass = (cpx * RGAS) / MW;
            if (k == ns) dy = qr / (rho * CpMixMass);
        }
        V[dst][0] = k < ne ? dy : 0.;
    }

    // Constant-volume reactor: density of the cell, or the ideal-gas density of the
    // initial state y0 (thread value) at the reactor pressure.
    __device__ __forceinline__ void set_density(double y0, double T0, const double *rho_cells, int64_t cell) {
        const double mwk = tid < ns ? L.sp[15 * L.SP + tid] : 1.;
        const double MW0 = 1. / wg_sum(tid < ns ? y0 / mwk : 0.);
        rho0 = rho_cells ? rho_cells[cell] : K.pressure * MW0 / (RGAS * T0);
    }

    // Strang "compact" sources of the final state (chemistry_DI.H:236-282); T_f is
    // passed in (thread ns holds it).
    __device__ __forceinline__ double compact_sources(double y, double yin, double Tf, double Tin, double dtc) {
        const bool isp = tid < ns;
        const double mwk = isp ? L.sp[15 * L.SP + tid] : 1.;
        double cp = 0., h, s, dcp;
        if (isp) thermo(tid, Tf, cp, h, s, dcp);
        double u = isp ? y / mwk : 0., cy = isp ? cp * (y / mwk) : 0.;
        wg_sum2(u, cy);
        const double rho = K.pressure * (1. / u) / (RGAS * Tf);
        const bool ok = dtc > 1.e-14;
        if (isp) return ok ? rho * (y - yin) / dtc : 0.;
        return (ok && K.energy) ? rho * (cy * RGAS) * (Tf - Tin) / dtc : 0.;
    }

    // KernelDense::NumericalJacobian, max-constraint branch (KernelDense.hpp:160-183)
    __device__ __forceinline__ void jacobian_fd(double h) {
        const double ZERO_DER = sqrt(TINY_F), ETA2 = sqrt(EPS_D), BETA = 1.e+3 * EPS_D;
        const double fw = V[F_][0] * V[OOE_][0];
        double hf = BETA * fabs(h) * sqrt(wg_sum(fw * fw) / (double)ne) * (double)ne;
        if (hf < 1.e-10) hf = 1.;
        for (int j = 0; j < ne; j++) {
            double *b = red + 8 * rp;
            rp ^= 1;
            if (tid == j) { b[0] = V[VB_][0]; b[1] = V[OOE_][0]; }
            __syncthreads();
            const double yh = b[0], ej = b[1];
            const double a = fabs(yh), bb = fabs(1. / ej);
            double hJ = ETA2 * (a > bb ? a : bb);
            hJ = fmax(hJ, hf / ej);
            hJ = fmax(hJ, ZERO_DER);
            const double ymx = j < ns ? 1. : 6000.;
            if (yh + hJ > ymx) hJ = -hJ;
            if (tid == j) V[VB_][0] = yh + hJ;
            rhs(VB_, AUX_);
            if (tid == j) V[VB_][0] = yh;
            const double hInv = 1. / hJ;
            if (tid < ne) J[(size_t)j * ne + tid] = hInv * (V[AUX_][0] - V[F_][0]);
        }
    }

    // Accumulator of the analytic Jacobian's reaction loop for one block of rows: dR_k/dc_i for
    // i in [i0, i1) in FP64 (D64, row stride ldd), R_k and dR_k/dT only in the first block.
    struct WideJacAcc {
        double *sRw, *sDTw;
        double *D64;
        int ldd, i0, i1, ns;
        bool first;
        __device__ __forceinline__ void R(unsigned k, double r, double drdT) const {
            if (!first) return;
            atomicAdd(&sRw[k], r);
            atomicAdd(&sDTw[k], drdT);
        }
        __device__ __forceinline__ void D(unsigned i, unsigned k, double v) const {
            if ((int)i >= i0 && (int)i < i1 && (int)k < ns) atomicAdd(&D64[(i - i0) * ldd + k], v);
        }
        __device__ __forceinline__ void U(unsigned k, double v) const {
            if ((int)k < ns)
                for (int i = i0; i < i1; i++) atomicAdd(&D64[(i - i0) * ldd + k], v);
        }
    };

    // Analytic Jacobian (same closed forms as the team kernel).  dR/dc is gathered in FP64 in
    // the LU area of LDS, which holds a block of its rows: the reaction loop runs once per block
    // (two blocks for the mechanisms here), each block's rows go to J as they are, and a last
    // sweep over J completes the entries that need all of dR/dc (S_k = sum_i dR_k/dc_i c_i).
    __device__ __forceinline__ void jacobian_analytic() {
        const int k = tid;
        const bool isp = k < ns;
        const double yv = V[VB_][0];
        const int P = L.SP;
        const double mwk = isp ? L.sp[15 * P + k] : 1.;
        double T;
        const double MW = 1. / wg_sum(isp ? yv / mwk : 0., &T, yv, ns);
        const double Yk = isp ? yv : 0.;
        const double xk = (MW * Yk) / mwk;
        const double cTotP = K.pressure / RGAS / T;
        const double ck = cTotP * xk;
        const double rho = cTotP * MW;
        const double uT = 1. / T, logT = log(T);
        double cp = 0., h = 0., s = 0., dcp = 0.;
        if (isp) {
            thermo(k, T, cp, h, s, dcp);
            scg[k] = make_double2(ck, h - s);
            sH[k] = h;
        }
        if (k == (int)dummy) { scg[k] = make_double2(1., 0.); sH[k] = 0.; }
#pragma unroll
        for (int w = 0; w < 4; w++) { sR[w * 256 + k] = 0.; sDT[w * 256 + k] = 0.; }
        double cTot = ck, cpx = isp ? cp * xk : 0.;
        wg_sum2(cTot, cpx);
        const double logPRT = log(101325. / RGAS / T);
        double *D64 = (double *)A;
        const int ldd = ns | 1;
        const int rows = (int)(((size_t)4 * (ne + 1) * lda) / ((size_t)8 * ldd));
        double S = 0., Acol = 0.;
        for (int i0 = 0; i0 < ns; i0 += rows) {
            const int i1 = i0 + rows < ns ? i0 + rows : ns;
            for (int idx = tid; idx < (i1 - i0) * ldd; idx += WIDE_THREADS) D64[idx] = 0.;
            __syncthreads();
            const WideJacAcc acc{sR + 256 * wave, sDT + 256 * wave, D64, ldd, i0, i1, ns, i0 == 0};
            // The waves scatter into the shared block one after another: FP64 sums are
            // order-dependent, and concurrent waves would make the Jacobian (and so the step
            // sequence) vary from run to run.  Within a wave the LDS atomics are applied in a
            // fixed order.
            for (int w = 0; w < WIDE_THREADS / 64; w++) {
                if (wave == w)
                    for (int j = tid; j < nr; j += WIDE_THREADS)
                        rxn_jac<XK>(L, scg, sH, ns, dummy, j, T, logT, uT, logPRT, K.pressure, cTot, acc);
                __syncthreads();
            }
            if (isp) {
                for (int i = i0; i < i1; i++) S += D64[(i - i0) * ldd + k] * scg[i].x;
                if (k >= i0 && k < i1)
                    for (int i = 0; i < ns; i++) Acol += sH[i] * D64[(k - i0) * ldd + i];
                for (int j = i0; j < i1; j++) J[(size_t)j * ne + k] = D64[(j - i0) * ldd + k];
            }
            __syncthreads();   // the block is overwritten by the next one
        }
        const double Rk = isp ? ((sR[k] + sR[256 + k]) + (sR[512 + k] + sR[768 + k])) : 0.;
        const double dRdT = isp ? ((sDT[k] + sDT[256 + k]) + (sDT[512 + k] + sDT[768 + k])) : 0.;
        const double fk = isp ? (mwk * Rk) / rho : 0.;
        if (isp) {
            for (int j = 0; j < ns; j++) {
                const double mwj = L.sp[15 * P + j];
                const double Djk = J[(size_t)j * ne + k];
                J[(size_t)j * ne + k] = (mwk / (rho * mwj)) * (rho * Djk - MW * S) + fk * MW / mwj;
            }
            J[(size_t)ns * ne + k] = (mwk / rho) * (dRdT - S * uT) + fk * uT;
        }
        const double cpm = (cpx * RGAS) / MW;
        const double Phi = rho * cpm;
        double hs = isp ? h * Rk : 0., hS = isp ? h * S : 0.;
        wg_sum2(hs, hS);
        const double QR = -hs * RGAS * T;
        const double fT = K.energy ? QR / Phi : 0.;
        double dQ = isp ? -(cp * RGAS * Rk + h * RGAS * T * (dRdT - S * uT)) : 0.;
        double dP = isp ? Yk * dcp * RGAS / mwk : 0.;
        wg_sum2(dQ, dP);
        if (K.energy) {
            if (isp) {
                const double dQdY = -(RGAS * T / mwk) * (rho * Acol - MW * hS);
                const double dPdY = (rho / mwk) * (cp * RGAS - MW * cpm);
                J[(size_t)k * ne + ns] = (dQdY - fT * dPdY) / Phi;
            }
            if (k == ns) J[(size_t)ns * ne + ns] = (dQ - fT * (-Phi * uT + rho * dP)) / Phi;
        } else {
            if (isp) J[(size_t)k * ne + ns] = 0.;
            if (k == ns) J[(size_t)ns * ne + ns] = 0.;
        }
        __syncthreads();   // A is reused by the factorisation; J columns complete
    }

    __device__ __forceinline__ int jacobian(double h) {
        LSM_TIC(t0);
#ifdef LSM_PHASE_TIMERS
        in_jac = 1;
#endif
        int nrhs = 0;
        if (K.jac_mode == 1) {
            jacobian_analytic();
        } else {
            jacobian_fd(h);
            nrhs = ne;
            __syncthreads();
        }
#ifdef LSM_PHASE_TIMERS
        in_jac = 0;
#endif
        LSM_TOC(1, t0);
        return nrhs;
    }

    // G = I - hr0 J and its inverse (the Newton matrix LU of KernelDense.hpp:196-221, done as a
    // Gauss-Jordan inversion with the same partial pivoting), in FP32 in LDS, 16 columns per panel:
    //  1. the panel's remaining rows, one per thread, are factorised in registers to choose the
    //     pivots (largest |entry| among rows not yet chosen, the lowest row on a tie): one
    //     barrier per column, each wave's candidate writing its row to a slot of its own before
    //     it, so that every thread reads the winner's row after it;
    //  2. wave 0 inverts the pivot block, T = A[pivot rows][panel]^-1, by Gauss-Jordan in
    //     registers (lane i: row i; its pivots are the ones chosen above, in order);
    //  3. every thread swaps the rows of its own column (the pivots as interchanges);
    //  4. X = T A[pivot rows][other columns]                  (MFMA 16x16x4, one tile per wave)
    //  5. A[other rows][other columns] -= A[other rows][panel] X     (MFMA, the rank-16 update)
    //  6. A[other rows][panel] = -A[other rows][panel] T, A[pivot rows][panel] = T;
    //     5 and 6 by 16-row blocks, each owned by one wave.
    // After the last panel A holds (P G)^-1, P the row interchanges (perm[i]: the row of G at
    // position i), so G^-1 b = A (P b): each FP32 solve is one mat-vec.  MFMA operands are loaded
    // from clamped addresses unconditionally and masked afterwards (a masked load would become a
    // branch with a wait of its own), all of a tile's loads issued before its first MFMA (a
    // scheduling barrier: otherwise the compiler moves MFMAs up between the loads and waits on each
    // load on its own); masked-off results are stored to a per-lane dummy slot instead of under a
    // branch.
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    __device__ __forceinline__ static f32x4 mfma16(float a, float b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // Position of the x-th row (column) outside the panel [k0, k0 + w).
    __device__ __forceinline__ static int skip_panel(int x, int k0, int w) { return x < k0 ? x : x + w; }

    // Wavefront maximum of floats >= -1 (lanes without a DPP source read 0, so a wave of -1
    // reports 0; callers check eligibility with a ballot).
    __device__ __forceinline__ static float wave_max_f(float x) {
        x = fmaxf(x, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x111, 0xf, 0xf, false)));
        x = fmaxf(x, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x112, 0xf, 0xf, false)));
        x = fmaxf(x, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x114, 0xf, 0xf, false)));
        x = fmaxf(x, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x118, 0xf, 0xf, false)));
        x = fmaxf(x, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x142, 0xa, 0xf, false)));
        x = fmaxf(x, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x143, 0xc, 0xf, false)));
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
    }

    // Step 1: the panel's pivot rows (positions before the interchanges) in prow[0..w); thread j
    // < w also keeps its own, prow[j], in mine.  Per column: each wave's largest |entry| among
    // its eligible rows (NaN counts as +inf, so some eligible row always wins) and the lowest such
    // row (a ballot), its row written to the wave's slot; after the barrier every thread takes
    // the largest over the waves, the lowest wave on a tie (rows grow with the wave index).
    __device__ __forceinline__ void panel_pivots(int k0, int w, int (&prow)[16], int &mine) {
        const int n = ne;
        float *slot = fsc + 800;                  // [2][4][16] candidate rows by parity and wave
        float *mslot = fsc + 928;                 // [2][4] wave maxima (-1: no eligible row)
        int *rslot = (int *)(fsc + 936);          // [2][4] their rows
        const int row = k0 + tid;
        bool elig = row < n;
        float pr[16];
        {
            const float *Ar = A + (elig ? row : 0) * lda + k0;
#pragma unroll
            for (int j = 0; j < 16; j++) pr[j] = Ar[j < w ? j : 0];
            asm volatile("" ::: "memory");
#pragma unroll
            for (int j = 0; j < 16; j++) pr[j] = (elig && j < w) ? pr[j] : 0.f;
        }
#pragma unroll
        for (int j = 0; j < 16; j++) {
            prow[j] = 0;
            if (j < w) {
                const int par = j & 1;
                const float av = fabsf(pr[j]);
                const float a = elig ? (av == av ? av : __builtin_huge_valf()) : -1.f;
                const float m = wave_max_f(a);
                const unsigned long long hit = __ballot(elig && a == m);
                const int wl = hit ? __ffsll(hit) - 1 : -1;          // the wave's lowest such lane
                if (lane == wl) {
                    float4 *dst = (float4 *)(slot + (par * 4 + wave) * 16);
                    dst[0] = make_float4(pr[0], pr[1], pr[2], pr[3]);
                    dst[1] = make_float4(pr[4], pr[5], pr[6], pr[7]);
                    dst[2] = make_float4(pr[8], pr[9], pr[10], pr[11]);
                    dst[3] = make_float4(pr[12], pr[13], pr[14], pr[15]);
                }
                if (lane == 0) {
                    mslot[par * 4 + wave] = hit ? m : -1.f;
                    rslot[par * 4 + wave] = k0 + 64 * wave + wl;
                }
                __syncthreads();
                const float4 mv = *(const float4 *)(mslot + par * 4);
                const int4 rv = *(const int4 *)(rslot + par * 4);
                float best = mv.x;
                int ww = 0, p = rv.x;
                if (mv.y > best) { best = mv.y; ww = 1; p = rv.y; }
                if (mv.z > best) { best = mv.z; ww = 2; p = rv.z; }
                if (mv.w > best) { best = mv.w; ww = 3; p = rv.w; }
                // uniform across the workgroup (every thread read the same slots)
                p = __builtin_amdgcn_readfirstlane(p);
                const int pw = __builtin_amdgcn_readfirstlane(ww);
                prow[j] = p;
                if (tid == j) mine = p;
                const float4 *src = (const float4 *)(slot + (par * 4 + pw) * 16);
                float u[16];
                {
                    const float4 u0 = src[0], u1 = src[1], u2 = src[2], u3 = src[3];
                    u[0] = u0.x; u[1] = u0.y; u[2] = u0.z; u[3] = u0.w;
                    u[4] = u1.x; u[5] = u1.y; u[6] = u1.z; u[7] = u1.w;
                    u[8] = u2.x; u[9] = u2.y; u[10] = u2.z; u[11] = u2.w;
                    u[12] = u3.x; u[13] = u3.y; u[14] = u3.z; u[15] = u3.w;
                }
                if (row == p) elig = false;
                const float rpiv = u[j] != 0.f ? 1.f / u[j] : 0.f;
                const float l = elig ? pr[j] * rpiv : 0.f;
#pragma unroll
                for (int jj = j + 1; jj < 16; jj++) pr[jj] = fmaf(-l, u[jj], pr[jj]);
            }
        }
    }

    // Step 2 (wave 0): T = A11^-1 by Gauss-Jordan without further pivoting, lane i < w holding
    // row i of A11 = A[prow][panel] (pi: prow[i]); rows and columns past w stay the identity.
    __device__ __forceinline__ void pivot_block_inverse(int k0, int w, int pi, float *Ts) {
        float t[16];
        const bool own = lane < w;
        const float *Ar = A + (own ? pi : 0) * lda + k0;
#pragma unroll
        for (int j = 0; j < 16; j++) t[j] = Ar[j < w ? j : 0];
        asm volatile("" ::: "memory");
#pragma unroll
        for (int j = 0; j < 16; j++) t[j] = (own && j < w) ? t[j] : (lane == j ? 1.f : 0.f);
#pragma unroll
        for (int j = 0; j < 16; j++) {
            if (j < w) {
                const float piv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t[j]), j));
                const float rp = piv != 0.f ? 1.f / piv : 0.f;
                const float f = t[j];
                if (lane == j) {
#pragma unroll
                    for (int jj = 0; jj < 16; jj++) t[jj] = jj == j ? rp : t[jj] * rp;
                }
                float rj[16];
#pragma unroll
                for (int jj = 0; jj < 16; jj++) rj[jj] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(t[jj]), j));
                if (lane != j) {
#pragma unroll
                    for (int jj = 0; jj < 16; jj++) t[jj] = jj == j ? -f * rp : fmaf(-f, rj[jj], t[jj]);
                }
            }
        }
        if (lane < 16) {
            float4 *dst = (float4 *)(Ts + lane * 16);
            dst[0] = make_float4(t[0], t[1], t[2], t[3]);
            dst[1] = make_float4(t[4], t[5], t[6], t[7]);
            dst[2] = make_float4(t[8], t[9], t[10], t[11]);
            dst[3] = make_float4(t[12], t[13], t[14], t[15]);
        }
    }

    __device__ __forceinline__ void factor_impl(double hr) {
        hr0 = hr;
        fp64 = false;
        const int n = ne;
        float *Ts = fsc + 512;                    // [16][16] inverse of the pivot block
        if (tid < n) {
            const size_t t = tid;
            int j = 0;
            for (; j + 4 <= n; j += 4) {
                const double j0 = J[(size_t)j * n + t], j1 = J[(size_t)(j + 1) * n + t];
                const double j2 = J[(size_t)(j + 2) * n + t], j3 = J[(size_t)(j + 3) * n + t];
                A[tid * lda + j] = (float)((tid == j ? 1. : 0.) - hr * j0);
                A[tid * lda + j + 1] = (float)((tid == j + 1 ? 1. : 0.) - hr * j1);
                A[tid * lda + j + 2] = (float)((tid == j + 2 ? 1. : 0.) - hr * j2);
                A[tid * lda + j + 3] = (float)((tid == j + 3 ? 1. : 0.) - hr * j3);
            }
            for (; j < n; j++) A[tid * lda + j] = (float)((tid == j ? 1. : 0.) - hr * J[(size_t)j * n + t]);
            perm[tid] = tid;
        }
        __syncthreads();
        const int il = lane & 15, kl = lane >> 4;     // MFMA operand coordinates of this lane
        for (int k0 = 0; k0 < n; k0 += 16) {
            const int w = n - k0 < 16 ? n - k0 : 16;
            int prow[16], mine = 0;
            panel_pivots(k0, w, prow, mine);
            if (wave == 0) pivot_block_inverse(k0, w, mine, Ts);
            // the pivots as interchanges: step j swaps positions k0 + j and ipiv[j]
            int ipiv[16];
#pragma unroll
            for (int j = 0; j < 16; j++) {
                ipiv[j] = prow[j];
#pragma unroll
                for (int jj = j + 1; jj < 16; jj++)
                    if (prow[jj] == k0 + j) prow[jj] = prow[j];   // the row at k0 + j moves to prow[j]
            }
            // 3. row interchanges, one column per thread (the permutation by the last thread);
            // the panel columns are read by the inverse above, so this waits for it
            __syncthreads();
            if (tid < n) {
#pragma unroll
                for (int j = 0; j < 16; j++) {
                    const int a = k0 + j, b = ipiv[j];
                    if (j < w && a != b) {
                        const float va = A[a * lda + tid], vb = A[b * lda + tid];
                        A[a * lda + tid] = vb;
                        A[b * lda + tid] = va;
                    }
                }
            } else if (tid == WIDE_THREADS - 1) {
#pragma unroll
                for (int j = 0; j < 16; j++) {
                    const int a = k0 + j, b = ipiv[j];
                    if (j < w) {
                        const int pa = perm[a];
                        perm[a] = perm[b];
                        perm[b] = pa;
                    }
                }
            }
            __syncthreads();
            const int no = n - w;                     // rows (columns) outside the panel
            const int nt = (no + 15) >> 4;
            // 4. X = T A[pivot rows][other columns], in place
            for (int tc = wave; tc < nt; tc += 4) {
                const int xc = 16 * tc + il;
                const bool okc = xc < no;
                const int col = skip_panel(okc ? xc : 0, k0, w);
                float av[4], bv[4];
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    const int kk = 4 * s + kl;
                    av[s] = Ts[il * 16 + kk];
                    bv[s] = A[(k0 + (kk < w ? kk : 0)) * lda + col];
                }
                asm volatile("" ::: "memory");
                __builtin_amdgcn_sched_barrier(0);     // every load issued before the first MFMA
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int s = 0; s < 4; s++) acc = mfma16(av[s], (okc && 4 * s + kl < w) ? bv[s] : 0.f, acc);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int i = 4 * kl + r;
                    float *dst = (okc && i < w) ? A + (k0 + i) * lda + col : fsc + 256 + lane;
                    *dst = acc[r];
                }
            }
            __syncthreads();
            // 5. and 6., by 16-row blocks of the other rows, each block owned by one wave:
            //    A[rows][Co] -= A[rows][panel] X, two column tiles at a time (their MFMA chains
            //    interleave and share the panel operand), then A[rows][panel] = -A[rows][panel] T
            //    (only this wave reads those entries, so no barrier between the two)
            for (int tr = wave; tr < nt; tr += 4) {
                const int xa = 16 * tr + il;
                const bool oka = xa < no;
                const int arow = skip_panel(oka ? xa : 0, k0, w);
                float av[4];
                int crow[4];
                bool okr[4];
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    const int kk = 4 * s + kl;
                    av[s] = A[arow * lda + k0 + (kk < w ? kk : 0)];
                }
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int xr = 16 * tr + 4 * kl + r;
                    okr[r] = xr < no;
                    crow[r] = skip_panel(okr[r] ? xr : 0, k0, w);
                }
                asm volatile("" ::: "memory");
                float na[4];
#pragma unroll
                for (int s = 0; s < 4; s++) na[s] = (oka && 4 * s + kl < w) ? -av[s] : 0.f;
                for (int tc = 0; tc < nt; tc += 2) {
                    const int xc0 = 16 * tc + il, xc1 = xc0 + 16;
                    const bool ok0 = xc0 < no, ok1 = xc1 < no;
                    const int col0 = skip_panel(ok0 ? xc0 : 0, k0, w), col1 = skip_panel(ok1 ? xc1 : 0, k0, w);
                    float b0[4], b1[4];
                    f32x4 c0, c1;
#pragma unroll
                    for (int s = 0; s < 4; s++) {
                        const int kk = 4 * s + kl;
                        const int prw = (k0 + (kk < w ? kk : 0)) * lda;
                        b0[s] = A[prw + col0];
                        b1[s] = A[prw + col1];
                    }
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        c0[r] = A[crow[r] * lda + col0];
                        c1[r] = A[crow[r] * lda + col1];
                    }
                    asm volatile("" ::: "memory");
                    __builtin_amdgcn_sched_barrier(0); // all 16 loads in flight before the MFMAs
#pragma unroll
                    for (int s = 0; s < 4; s++) {
                        const bool okk = 4 * s + kl < w;
                        c0 = mfma16(na[s], (ok0 && okk) ? b0[s] : 0.f, c0);
                        c1 = mfma16(na[s], (ok1 && okk) ? b1[s] : 0.f, c1);
                    }
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        float *d0 = (okr[r] && ok0) ? A + crow[r] * lda + col0 : fsc + 256 + lane;
                        float *d1 = (okr[r] && ok1) ? A + crow[r] * lda + col1 : fsc + 320 + lane;
                        *d0 = c0[r];
                        *d1 = c1[r];
                    }
                }
                float tv[4];
#pragma unroll
                for (int s = 0; s < 4; s++) tv[s] = Ts[(4 * s + kl) * 16 + il];
                __builtin_amdgcn_sched_barrier(0);
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int s = 0; s < 4; s++) acc = mfma16((oka && 4 * s + kl < w) ? av[s] : 0.f, tv[s], acc);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    float *dst = (okr[r] && il < w) ? A + crow[r] * lda + k0 + il : fsc + 256 + lane;
                    *dst = -acc[r];
                }
            }
            {
                const int i = tid >> 4, j = tid & 15;
                if (i < w && j < w) A[(k0 + i) * lda + k0 + j] = Ts[i * 16 + j];
            }
            __syncthreads();
        }
    }
    __device__ __forceinline__ void factor(double hr) { LSM_TIC(t0); factor_impl(hr); LSM_TOC(2, t0); }

    // bs <- G^-1 bs (FP64 in, FP64 out) in FP32: x = A (P b), A = (P G)^-1 from factor_impl, one
    // row per thread, four partial sums.  Accuracy comes from the FP64 residuals of iterative
    // refinement.
    __device__ __forceinline__ void lu_solve_bs() {
        const int n = ne;
        float *ys = fsc;                              // FP32 (P b) [256]
        if (tid < n) ys[tid] = (float)bs[perm[tid]];
        __syncthreads();
        const int row = tid < n ? tid : n - 1;
        const float *Ar = A + row * lda;
        float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
        int j = 0;
        for (; j + 16 <= n; j += 16) {
            float v[16], x[16];
#pragma unroll
            for (int u = 0; u < 16; u++) { v[u] = Ar[j + u]; x[u] = ys[j + u]; }
#pragma unroll
            for (int u = 0; u < 16; u += 4) {
                a0 = fmaf(v[u], x[u], a0);
                a1 = fmaf(v[u + 1], x[u + 1], a1);
                a2 = fmaf(v[u + 2], x[u + 2], a2);
                a3 = fmaf(v[u + 3], x[u + 3], a3);
            }
        }
        for (; j < n; j++) a0 = fmaf(Ar[j], ys[j], a0);
        const float x = (a0 + a1) + (a2 + a3);
        if (tid < n) bs[tid] = (double)x;
        __syncthreads();
    }

    // Row `tid` of J times the vector in bs (loads unconditional, 4 partial sums: column j goes to
    // a[j % 4] below the largest multiple of 4 not above n, the rest to a0).  The Jacobian sits in
    // the workgroup's HBM workspace: its loads are issued 32 at a time, so that the latency of one
    // batch is paid once, not per column.
    __device__ __forceinline__ double jrow_dot() const {
        const int n = ne;
        const int row = tid < n ? tid : n - 1;
        const double *Jr = J + row;
        double a0 = 0., a1 = 0., a2 = 0., a3 = 0.;
        int j = 0;
        for (; j + 32 <= n; j += 32) {
            double v[32];
#pragma unroll
            for (int u = 0; u < 32; u++) v[u] = Jr[(size_t)(j + u) * n];
#pragma unroll
            for (int u = 0; u < 32; u += 4) {
                a0 = fma(v[u], bs[j + u], a0);
                a1 = fma(v[u + 1], bs[j + u + 1], a1);
                a2 = fma(v[u + 2], bs[j + u + 2], a2);
                a3 = fma(v[u + 3], bs[j + u + 3], a3);
            }
        }
        for (; j + 4 <= n; j += 4) {
            a0 = fma(J[(size_t)j * n + row], bs[j], a0);
            a1 = fma(J[(size_t)(j + 1) * n + row], bs[j + 1], a1);
            a2 = fma(J[(size_t)(j + 2) * n + row], bs[j + 2], a2);
            a3 = fma(J[(size_t)(j + 3) * n + row], bs[j + 3], a3);
        }
        for (; j < n; j++) a0 = fma(J[(size_t)j * n + row], bs[j], a0);
        return (a0 + a1) + (a2 + a3);
    }

    // FP64 LU of G = I - hr0 J with partial pivoting and physical row interchanges (the textbook
    // form of KernelDense.hpp:196-221), column-major in the HBM workspace G64; thread i owns row
    // i.  The fallback of a factorisation whose FP32 solves do not refine to the FP64 accuracy.
    __device__ void factor64() {
        const int n = ne;
        for (int e = tid; e < n * n; e += WIDE_THREADS) {
            const int i = e % n, j = e / n;
            G64[e] = (i == j ? 1. : 0.) - hr0 * J[e];
        }
        __syncthreads();
        for (int k = 0; k < n; k++) {
            const double a = (tid < n && tid >= k) ? fabs(G64[(size_t)k * n + tid]) : -1.;
            // largest |a|, lowest row on a tie
            unsigned long long key = a >= 0. ? (((unsigned long long)__double_as_longlong(a) & ~0xffull) |
                                                (unsigned long long)(255 - tid)) : 0ull;
            key = wave_max_u64(key);
            unsigned long long *kb = (unsigned long long *)(red + 8 * rp);
            rp ^= 1;
            if (lane == 0) kb[wave] = key;
            __syncthreads();
            key = umax64(umax64(kb[0], kb[1]), umax64(kb[2], kb[3]));
            const int p = key == 0ull ? k : 255 - (int)(key & 0xffull);
            if (tid == 0) p64[k] = p;
            if (p != k)
                for (int j = tid; j < n; j += WIDE_THREADS) {
                    const double t = G64[(size_t)j * n + k];
                    G64[(size_t)j * n + k] = G64[(size_t)j * n + p];
                    G64[(size_t)j * n + p] = t;
                }
            __syncthreads();
            const double piv = G64[(size_t)k * n + k];
            double l = 0.;
            if (tid > k && tid < n && piv != 0.) {
                l = G64[(size_t)k * n + tid] / piv;
                G64[(size_t)k * n + tid] = l;
            }
            if (tid > k && tid < n)
                for (int j = k + 1; j < n; j++) G64[(size_t)j * n + tid] = fma(-l, G64[(size_t)j * n + k], G64[(size_t)j * n + tid]);
            __syncthreads();
        }
        fp64 = true;
    }
    // bs <- G^-1 bs with the FP64 factors (row interchanges, then L and U by columns).
    __device__ void solve64_bs() {
        const int n = ne;
        if (tid == 0)
            for (int k = 0; k < n; k++) {
                const int p = p64[k];
                if (p != k) { const double t = bs[k]; bs[k] = bs[p]; bs[p] = t; }
            }
        __syncthreads();
        for (int k = 0; k < n; k++) {
            const double xk = bs[k];
            if (tid > k && tid < n) bs[tid] = fma(-G64[(size_t)k * n + tid], xk, bs[tid]);
            __syncthreads();
        }
        for (int k = n - 1; k >= 0; k--) {
            if (tid == k) bs[k] = bs[k] / G64[(size_t)k * n + k];
            __syncthreads();
            const double xk = bs[k];
            if (tid < k) bs[tid] = fma(-G64[(size_t)k * n + tid], xk, bs[tid]);
            __syncthreads();
        }
    }

    // G^-1 b: the FP32 solve, then up to K.refine sweeps of iterative refinement with the residual
    // in FP64 from the FP64 Jacobian, r = b - (x - hr0 J x), until the correction's WRMS norm in
    // the integration's own weights is below REFINE_TOL (1e-5 of the tolerance, far below the Newton
    // test l0 ||dx|| < 1).  When a sweep fails to halve the correction, or the sweeps run out, the
    // factorisation switches to FP64 (factor64) and this and the following solves use it.
    // K.refine = 0: the FP32 solve alone, unchecked.
    static constexpr double REFINE_TOL = 1.e-5;
    __device__ __forceinline__ void solve_impl(int vid) {
        const double b = V[vid][0];
        if (tid < ne) bs[tid] = b;
        __syncthreads();
        if (fp64) {
            solve64_bs();
            V[vid][0] = tid < ne ? bs[tid] : 0.;
            __syncthreads();
            return;
        }
        LSM_TIC(ts0);
        lu_solve_bs();
        LSM_TOC(7, ts0);
        double x = tid < ne ? bs[tid] : 0.;
        bool ok = K.refine == 0;
        double prev = BIG_D;
        for (int it = 0; it < K.refine; it++) {
            LSM_TIC(tr0);
            const double jx = jrow_dot();
            LSM_TOC(8, tr0);
#ifdef LSM_PHASE_TIMERS
            tp[9] += 1;
#endif
            const double r = b - (x - hr0 * jx);
            __syncthreads();
            if (tid < ne) bs[tid] = r;
            __syncthreads();
            LSM_TIC(ts1);
            lu_solve_bs();
            LSM_TOC(7, ts1);
            const double d = tid < ne ? bs[tid] : 0.;
            x += d;
            const double dw = d * V[OOE_][0];
            const double nd = sqrt(wg_sum(dw * dw) / (double)ne);
            if (tid < ne) bs[tid] = x;
            __syncthreads();
            if (nd <= REFINE_TOL) { ok = true; break; }
            if (!(nd <= 0.5 * prev)) break;
            prev = nd;
        }
        if (!ok) {
            if (tid == 0 && K.fp64_fallbacks) atomicAdd(K.fp64_fallbacks, 1ull);
            factor64();
            if (tid < ne) bs[tid] = b;
            __syncthreads();
            solve64_bs();
            x = tid < ne ? bs[tid] : 0.;
            __syncthreads();
        }
        V[vid][0] = tid < ne ? x : 0.;
    }
    __device__ __forceinline__ void solve(int vid) { LSM_TIC(t0); solve_impl(vid); LSM_TOC(3, t0); }

    __device__ __forceinline__ void jac_times(int src, int dst) {
        LSM_TIC(t0);
        if (tid < ne) bs[tid] = V[src][0];
        __syncthreads();
        const double acc = jrow_dot();
        __syncthreads();
        V[dst][0] = tid < ne ? acc : 0.;
        LSM_TOC(4, t0);
    }
};

// One cell per workgroup, persistent over an atomic queue; same harness as the
// team kernel (chemistry_DI.H:92-328).
template <bool XK>
__global__ void __launch_bounds__(256) wide_kernel(DevMech M, KParams K, int64_t n, const double *__restrict__ dt,
                                                   const double *yin, double *yout, int *status, double *q,
                                                   int *counters, unsigned *queue, const int *order,
                                                   double *jws) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    WideCtx<XK> c(M, K, lds_raw, jws + (size_t)blockIdx.x * 2 * M.ne * M.ne);
    const int tid = threadIdx.x, ns = M.ns, ne = M.ne;
    int *cell_slot = (int *)(c.red + 16);
    for (;;) {
        if (tid == 0) {
            const unsigned t = atomicAdd(queue, 1u);
            cell_slot[0] = (int64_t)t < n ? 1 : 0;
            cell_slot[1] = (int64_t)t < n ? (order ? order[t] : (int)t) : -1;
        }
        __syncthreads();
        const int more = cell_slot[0];
        const int64_t cell = (int64_t)cell_slot[1];
        __syncthreads();
        if (!more) break;
        if (cell < 0 || cell >= n) continue;      // visiting order entries outside [0, n) are skipped
        LSM_TIC(tcell);
        const double yi = tid < ne ? yin[(int64_t)tid * n + cell] : 0.;
        double T0;
        c.wg_sum(0., &T0, yi, ns);
        double y;
        int st = 0;
        BdfCounters cnt = BdfCounters{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (T0 > K.min_temperature) {                                  // chemistry_DI.H:97
            double y0 = yi;                                            // :101-128
            if (tid < ns && y0 < 0.) y0 = 0.;
            const double sum = c.wg_sum(tid < ns ? y0 : 0.);
            if (tid < ns) y0 /= sum;
            c.V[Y0_][0] = tid < ne ? y0 : 0.;
            if (K.const_volume) c.set_density(y0, T0, K.rho, cell);     // :173
            st = bdf_integrate(c, K.bk, K.bk.t0 + dt[cell], cnt);       // :138-157
            y = st > 0 ? c.V[Y_][0] : c.V[Y0_][0];
            if (st > 0 && q != nullptr && tid == 0) q[cell] = c.qr;     // :168
        } else {
            y = yi;
        }
        double Tf;
        const double sum = c.wg_sum(tid < ns ? y : 0., &Tf, y, ns);   // :219 normalizeMassFractions
        if (tid < ns) y /= sum;
        if (!K.energy) { if (tid == ns) y = yi; Tf = T0; }
        if (K.compact) y = c.compact_sources(y, yi, Tf, T0, dt[cell]); // :236-282
        if (tid < ne) yout[(int64_t)tid * n + cell] = y;
        if (tid == 0 && status) status[cell] = st;
#ifdef LSM_PHASE_TIMERS
        c.tp[5] += clock64() - tcell;
        c.tp[6] += 1;
#endif
        if (counters != nullptr && tid < LSM_NCOUNTERS_DEV) {
            const unsigned vals[12] = {cnt.steps, cnt.fcalls, cnt.fcalls_jac, cnt.jacobians, cnt.factorizations,
                                       cnt.solves, cnt.conv_failures, cnt.err_failures, cnt.max_order,
                                       cnt.rhs_total, cnt.clamps, cnt.newton};
            unsigned v = 0;
#pragma unroll
            for (int i = 0; i < 12; i++)
                if (tid == i) v = vals[i];
            counters[cell * LSM_NCOUNTERS_DEV + tid] = (int)v;
        }
    }
#ifdef LSM_PHASE_TIMERS
    if (tid == 0 && K.phase_cycles)
        for (int i = 0; i < 12; i++) atomicAdd(&K.phase_cycles[i], c.tp[i]);
#endif
}

// Right-hand side and analytic Jacobian of the constant-pressure reactor at given
// states (GetEquations of the batch-reactor ODE system), one wave per cell, in
// grid-stride order over the cells so that the per-wave workspace of the
// integrator is reused.  y: SoA [ne][n]; f: SoA [ne][n]; J: [n][ne][ne] row-major.
template <bool XK>
__global__ void __launch_bounds__(256) equations_kernel(DevMech M, KParams K, int64_t n, const double *yin,
                                                        double *f, double *J) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    {
        const int4 *src = (const int4 *)M.img.src;
        int4 *dst = (int4 *)lds_raw;
        const int nv = M.img.bytes / 16;
        for (int i = threadIdx.x; i < nv; i += blockDim.x) dst[i] = src[i];
        __syncthreads();
    }
    const int wave = threadIdx.x >> 6, wpb = blockDim.x >> 6;
    double *wave_lds = (double *)(lds_raw + M.img.bytes + (size_t)wave * team_wave_bytes(M.ne));
    const size_t gw = (size_t)blockIdx.x * wpb + wave;
    TeamCtx<64, XK> c(M, K, lds_raw, wave_lds);
    const int lane = c.lane, ne = M.ne;
    for (int64_t cell = (int64_t)gw; cell < n; cell += (int64_t)gridDim.x * wpb) {
        c.V[VB_][0] = lane < ne ? yin[(int64_t)lane * n + cell] : 0.;
        if (K.const_volume) c.set_density(c.V[VB_][0], readlane_d(c.V[VB_][0], M.ns), K.rho, cell);
        c.rhs(VB_, F_);
        if (f != nullptr && lane < ne) f[(int64_t)lane * n + cell] = c.V[F_][0];
        if (J != nullptr) {
            c.jacobian_analytic_impl();
            double *Jc = J + cell * (int64_t)ne * ne;
            if (lane < ne)
                for (int j = 0; j < ne; j++) Jc[(int64_t)lane * ne + j] = c.Jl[j * 65 + lane];
        }
        wave_sync();
    }
}

// ---------------------------------------------------------------------------
// Plain context: one cell per thread, all state in HBM
// ---------------------------------------------------------------------------
template <bool XK>
struct PlainCtx {
    double *W;        // vectors  [NVEC][ne][slots]
    double *Jm, *Gm;  // matrices [ne][ne][slots]
    int *piv;         // [2][ne][slots] row / column permutation
    double *sc, *sh, *ss, *sr;
    const DevMech &M;
    const KParams &K;
    int64_t S;        // stride (slots)
    int64_t slot;
    int ne, ns, nr;
    double qr;
    double rho0;      // constant-volume reactor: density of the current cell
    __device__ __forceinline__ void add_time(int, unsigned long long) {}

    __device__ PlainCtx(const DevMech &m, const KParams &k, double *work, int64_t slots, int64_t sl)
        : M(m), K(k), S(slots), slot(sl), ne(m.ne), ns(m.ns), nr(m.nr), qr(0.), rho0(0.) {
        W = work;
        Jm = W + (int64_t)NVEC * ne * S;
        Gm = Jm + (int64_t)ne * ne * S;
        sc = Gm + (int64_t)ne * ne * S;
        sh = sc + (int64_t)ns * S;
        ss = sh + (int64_t)ns * S;
        sr = ss + (int64_t)ns * S;
        piv = (int *)(sr + (int64_t)nr * S);
    }

    template <class F>
    __device__ __forceinline__ void each(F f) {
        for (int e = 0; e < ne; e++) f(e);
    }
    __device__ __forceinline__ double &x(int vid, int e) { return W[((int64_t)vid * ne + e) * S + slot]; }
    __device__ __forceinline__ double ymin(int e) const { return e < ns ? 0. : 200.; }
    __device__ __forceinline__ double ymax(int e) const { return e < ns ? 1. : 6000.; }
    __device__ __forceinline__ double reduce(double v) { return v; }
    __device__ __forceinline__ int count(bool b) { return b ? 1 : 0; }
    __device__ __forceinline__ double &J(int i, int j) { return Jm[((int64_t)i * ne + j) * S + slot]; }
    __device__ __forceinline__ double &Gv(int i, int j) { return Gm[((int64_t)i * ne + j) * S + slot]; }
    __device__ __forceinline__ int &P(int w, int i) { return piv[((int64_t)w * ne + i) * S + slot]; }

    __device__ void rhs(int src, int dst) {
        const double T = x(src, ns);
        double u = 0.;
        for (int k = 0; k < ns; k++) u += x(src, k) / M.sp[SP_MW * M.ns_pad + k];
        const double MW = 1. / u;
        const double cTotP = K.const_volume ? rho0 / MW : K.pressure / RGAS / T;
        const double rho = K.const_volume ? rho0 : cTotP * MW;
        const double T2 = T * T, T3 = T2 * T, T4 = T3 * T, uT = 1. / T, logT = log(T);
        double cTot = 0., cpx = 0.;
        for (int k = 0; k < ns; k++) {
            const double xk = (MW * x(src, k)) / M.sp[SP_MW * M.ns_pad + k];
            const double ck = cTotP * xk;
            double cp, h, s;
            nasa7(M, k, T, T2, T3, T4, uT, logT, cp, h, s);
            sc[k * S + slot] = ck;
            sh[k * S + slot] = h;
            ss[k * S + slot] = s;
            cTot += ck;
            cpx += cp * xk;
        }
        const double logPRT = log(101325. / RGAS / T);
        const double Pk = K.const_volume ? cTotP * RGAS * T : K.pressure;   // the kinetics' pressure
        for (int j = 0; j < nr; j++)
            sr[(int64_t)j * S + slot] = reaction_rate<XK>(
                M, j, T, logT, uT, logPRT, Pk, cTot, [&](int i) { return sc[(int64_t)i * S + slot]; },
                [&](int i) { return sh[(int64_t)i * S + slot]; }, [&](int i) { return ss[(int64_t)i * S + slot]; });
        double hs = 0., rs = 0.;
        for (int k = 0; k < ns; k++) {
            double Rk = 0.;
            for (int e = M.f_ptr[k]; e < M.f_ptr[k + 1]; e++) Rk += M.f_nu[e] * sr[(int64_t)M.f_rx[e] * S + slot];
            x(dst, k) = (M.sp[SP_MW * M.ns_pad + k] * Rk) / rho;
            hs += Rk * sh[k * S + slot];
            rs += Rk;
        }
        double dT = 0.;
        if (K.energy && K.const_volume) {      // BatchReactorHomogeneousConstantVolume.H:185-196
            qr = -hs * RGAS * T;
            dT = (qr + RGAS * T * rs) / (rho * ((cpx * RGAS - RGAS) / MW));
        } else if (K.energy) {
            qr = -hs * RGAS * T;
            dT = qr / (rho * ((cpx * RGAS) / MW));
        }
        x(dst, ns) = dT;
    }

    __device__ int jacobian(double h) {
        const double ZERO_DER = sqrt(TINY_F), ETA2 = sqrt(EPS_D), BETA = 1.e+3 * EPS_D;
        double acc = 0.;
        for (int e = 0; e < ne; e++) { const double t = x(F_, e) * x(OOE_, e); acc += t * t; }
        double hf = BETA * fabs(h) * sqrt(acc / (double)ne) * (double)ne;
        if (hf < 1.e-10) hf = 1.;
        for (int j = 0; j < ne; j++) {
            const double yh = x(VB_, j);
            const double ej = x(OOE_, j);
            const double a = fabs(yh), b = fabs(1. / ej);
            double hJ = ETA2 * (a > b ? a : b);
            hJ = fmax(hJ, hf / ej);
            hJ = fmax(hJ, ZERO_DER);
            if (yh + hJ > ymax(j)) hJ = -hJ;
            x(VB_, j) = yh + hJ;
            rhs(VB_, AUX_);
            x(VB_, j) = yh;
            const double hInv = 1. / hJ;
            for (int i = 0; i < ne; i++) J(i, j) = hInv * (x(AUX_, i) - x(F_, i));
        }
        return ne;
    }

    // Eigen PartialPivLU / FullPivLU restated (KernelDense.hpp:196-221).
    __device__ void factor(double hr0) {
        for (int i = 0; i < ne; i++)
            for (int j = 0; j < ne; j++) Gv(i, j) = J(i, j) * -hr0 + (i == j ? 1. : 0.);
        for (int i = 0; i < ne; i++) { P(0, i) = i; P(1, i) = i; }
        const bool full = K.full_pivoting;
        for (int k = 0; k < ne; k++) {
            int p = k, qc = k;
            double best = -1.;
            if (!full) {
                for (int i = k; i < ne; i++) { const double a = fabs(Gv(i, k)); if (a > best) { best = a; p = i; } }
            } else {
                for (int j = k; j < ne; j++)
                    for (int i = k; i < ne; i++) { const double a = fabs(Gv(i, j)); if (a > best) { best = a; p = i; qc = j; } }
            }
            if (p != k) {
                for (int j = 0; j < ne; j++) { const double t = Gv(k, j); Gv(k, j) = Gv(p, j); Gv(p, j) = t; }
                const int t = P(0, k); P(0, k) = P(0, p); P(0, p) = t;
            }
            if (full && qc != k) {
                for (int i = 0; i < ne; i++) { const double t = Gv(i, k); Gv(i, k) = Gv(i, qc); Gv(i, qc) = t; }
                const int t = P(1, k); P(1, k) = P(1, qc); P(1, qc) = t;
            }
            const double pv = Gv(k, k);
            if (pv == 0.) continue;
            for (int i = k + 1; i < ne; i++) {
                const double l = Gv(i, k) / pv;
                Gv(i, k) = l;
                for (int j = k + 1; j < ne; j++) Gv(i, j) -= l * Gv(k, j);
            }
        }
    }

    __device__ void solve(int vid) {
        // t = P b (AUX as scratch), L t = t, U t = t, b = Q t
        for (int i = 0; i < ne; i++) x(AUX_, i) = x(vid, P(0, i));
        for (int i = 0; i < ne; i++) {
            double t = x(AUX_, i);
            for (int k = 0; k < i; k++) t -= Gv(i, k) * x(AUX_, k);
            x(AUX_, i) = t;
        }
        for (int i = ne - 1; i >= 0; i--) {
            double t = x(AUX_, i);
            for (int k = i + 1; k < ne; k++) t -= Gv(i, k) * x(AUX_, k);
            x(AUX_, i) = t / Gv(i, i);
        }
        if (K.full_pivoting) for (int i = 0; i < ne; i++) x(vid, P(1, i)) = x(AUX_, i);
        else for (int i = 0; i < ne; i++) x(vid, i) = x(AUX_, i);
    }

    __device__ void jac_times(int src, int dst) {
        for (int i = 0; i < ne; i++) {
            double acc = 0.;
            for (int j = 0; j < ne; j++) acc += J(i, j) * x(src, j);
            x(dst, i) = acc;
        }
    }
};

template <bool XK>
__global__ void __launch_bounds__(256) plain_kernel(DevMech M, KParams K, int64_t n, const double *__restrict__ dt,
                                                    const double *yin, double *yout, int *status, double *q,
                                                    int *counters, double *work, int64_t slots) {
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= slots) return;
    PlainCtx<XK> c(M, K, work, slots, slot);
    const int ns = M.ns, ne = M.ne;
    for (int64_t cell = slot; cell < n; cell += slots) {
        const double T0 = yin[(int64_t)ns * n + cell];
        int st = 0;
        BdfCounters cnt = BdfCounters{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (T0 > K.min_temperature) {
            double sum = 0.;
            for (int i = 0; i < ns; i++) {
                const double v = yin[(int64_t)i * n + cell];
                c.x(Y0_, i) = v < 0. ? 0. : v;
                sum += c.x(Y0_, i);
            }
            for (int i = 0; i < ns; i++) c.x(Y0_, i) /= sum;
            c.x(Y0_, ns) = T0;
            if (K.const_volume) {                                       // chemistry_DI.H:173
                double u = 0.;
                for (int i = 0; i < ns; i++) u += c.x(Y0_, i) / M.sp[SP_MW * M.ns_pad + i];
                c.rho0 = K.rho ? K.rho[cell] : K.pressure * (1. / u) / (RGAS * T0);
            }
            st = bdf_integrate(c, K.bk, K.bk.t0 + dt[cell], cnt);
            const int src = st > 0 ? Y_ : Y0_;
            for (int i = 0; i < ne; i++) c.x(AUX_, i) = c.x(src, i);
            if (st > 0 && q != nullptr) q[cell] = c.qr;
        } else {
            for (int i = 0; i < ne; i++) c.x(AUX_, i) = yin[(int64_t)i * n + cell];
        }
        double sum = 0.;
        for (int i = 0; i < ns; i++) sum += c.x(AUX_, i);
        for (int i = 0; i < ns; i++) c.x(AUX_, i) /= sum;
        if (!K.energy) c.x(AUX_, ns) = T0;
        if (K.compact) {                       // Strang "compact" sources, chemistry_DI.H:236-282
            const double Tf = c.x(AUX_, ns);
            const double T2 = Tf * Tf, T3 = T2 * Tf, T4 = T3 * Tf, uT = 1. / Tf, logT = log(Tf);
            double u = 0., cy = 0.;
            for (int i = 0; i < ns; i++) {
                double cp, h, s;
                nasa7(M, i, Tf, T2, T3, T4, uT, logT, cp, h, s);
                const double yw = c.x(AUX_, i) / M.sp[SP_MW * M.ns_pad + i];
                u += yw;
                cy += cp * yw;
            }
            const double rho = K.pressure * (1. / u) / (RGAS * Tf);
            const double dtc = dt[cell];
            const bool ok = dtc > 1.e-14;
            for (int i = 0; i < ns; i++) {
                const double yi = yin[(int64_t)i * n + cell];
                yout[(int64_t)i * n + cell] = ok ? rho * (c.x(AUX_, i) - yi) / dtc : 0.;
            }
            yout[(int64_t)ns * n + cell] = (ok && K.energy) ? rho * (cy * RGAS) * (Tf - T0) / dtc : 0.;
        } else {
            for (int i = 0; i < ne; i++) yout[(int64_t)i * n + cell] = c.x(AUX_, i);
        }
        if (status) status[cell] = st;
        if (counters) {
            int *o = counters + cell * LSM_NCOUNTERS_DEV;
            o[0] = cnt.steps; o[1] = cnt.fcalls; o[2] = cnt.fcalls_jac; o[3] = cnt.jacobians;
            o[4] = cnt.factorizations; o[5] = cnt.solves; o[6] = cnt.conv_failures; o[7] = cnt.err_failures;
            o[8] = cnt.max_order; o[9] = cnt.rhs_total; o[10] = cnt.clamps; o[11] = cnt.newton;
        }
    }
}

#if LSM_IN_PART(6)
__global__ void reset_queue_kernel(unsigned *queue) { *queue = 0u; }

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
int team_max_neq() { return 64; }

int team_waves_per_cu(const DevMech &m) { return 4 * (m.ne <= 16 ? TEAM_WAVES_PER_SIMD(16) : TEAM_WAVES_PER_SIMD(64)); }

bool team_shape_fits(const DevMech &m, int wpb, int blocks_per_cu, size_t lds_cu) {
    if (!m.img.team_ok || m.ne > 64 || wpb < 1 || wpb > 4 || blocks_per_cu < 1) return false;
    return (size_t)blocks_per_cu * ((size_t)m.img.bytes + (size_t)wpb * team_wave_bytes(m.ne)) <= lds_cu;
}

#endif  // LSM_IN_PART(6)

template <int NMAX, bool XK>
static int launch_team_n(const DevMech &m, const KParams &k, int64_t n, const double *dt, const double *yin,
                         double *yout, int *status, double *q, int *counters, unsigned *queue, const int *order,
                         int grid, int wpb, hipStream_t s) {
    const size_t lds = (size_t)m.img.bytes + (size_t)wpb * team_wave_bytes(m.ne);
    static size_t lds_set = 65536;   // largest dynamic LDS size granted so far (not re-set under graph capture)
    if (lds > lds_set) {
        if (hipFuncSetAttribute((const void *)team_kernel<NMAX, XK>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds) != hipSuccess)
            return -1;
        lds_set = lds;
    }
    hipLaunchKernelGGL((team_kernel<NMAX, XK>), dim3(grid), dim3(64 * wpb), lds, s, m, k, n, dt, yin, yout, status, q,
                       counters, queue, order);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// One launcher per register-array size, each in its own translation unit.
#define LSM_TEAM_LAUNCHER(NM)                                                                                       \
    int launch_team_##NM(const DevMech &m, const KParams &k, int64_t n, const double *dt, const double *yin,        \
                         double *yout, int *status, double *q, int *counters, unsigned *queue, const int *order,    \
                         int grid, int wpb, hipStream_t s)
#define LSM_TEAM_LAUNCHER_BODY(NM)                                                                                  \
    LSM_TEAM_LAUNCHER(NM) {                                                                                         \
        return m.xk ? launch_team_n<NM, true>(m, k, n, dt, yin, yout, status, q, counters, queue, order, grid, wpb, \
                                              s)                                                                    \
                    : launch_team_n<NM, false>(m, k, n, dt, yin, yout, status, q, counters, queue, order, grid,    \
                                               wpb, s);                                                             \
    }
LSM_TEAM_LAUNCHER(16);
LSM_TEAM_LAUNCHER(32);
LSM_TEAM_LAUNCHER(48);
LSM_TEAM_LAUNCHER(56);
LSM_TEAM_LAUNCHER(64);
#if LSM_IN_PART(1)
LSM_TEAM_LAUNCHER_BODY(16)
#endif
#if LSM_IN_PART(2)
LSM_TEAM_LAUNCHER_BODY(32)
#endif
#if LSM_IN_PART(3)
LSM_TEAM_LAUNCHER_BODY(48)
#endif
#if LSM_IN_PART(4)
LSM_TEAM_LAUNCHER_BODY(56)
#endif
#if LSM_IN_PART(5)
LSM_TEAM_LAUNCHER_BODY(64)
#endif

template <int SEG, bool XK>
static int launch_pack_n(const DevMech &m, const KParams &k, int64_t n, const double *dt, const double *yin,
                         double *yout, int *status, double *q, int *counters, unsigned *queue, const int *order,
                         int grid, int wpb, hipStream_t s) {
    const size_t lds = (size_t)m.img.bytes + (size_t)wpb * pack_wave_bytes(SEG, m.ne);
    static size_t lds_set = 65536;   // largest dynamic LDS size granted so far (not re-set under graph capture)
    if (lds > lds_set) {
        if (hipFuncSetAttribute((const void *)pack_kernel<SEG, XK>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds) != hipSuccess)
            return -1;
        lds_set = lds;
    }
    hipLaunchKernelGGL((pack_kernel<SEG, XK>), dim3(grid), dim3(64 * wpb), lds, s, m, k, n, dt, yin, yout, status, q,
                       counters, queue, order);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Packed kernels (2 or 4 cells per wavefront), one translation unit per segment width.
#define LSM_PACK_LAUNCHER(SG)                                                                                       \
    int launch_pack_##SG(const DevMech &m, const KParams &k, int64_t n, const double *dt, const double *yin,        \
                         double *yout, int *status, double *q, int *counters, unsigned *queue, const int *order,    \
                         int grid, int wpb, hipStream_t s)
#define LSM_PACK_LAUNCHER_BODY(SG)                                                                                  \
    LSM_PACK_LAUNCHER(SG) {                                                                                         \
        return m.xk ? launch_pack_n<SG, true>(m, k, n, dt, yin, yout, status, q, counters, queue, order, grid, wpb, \
                                              s)                                                                    \
                    : launch_pack_n<SG, false>(m, k, n, dt, yin, yout, status, q, counters, queue, order, grid,    \
                                               wpb, s);                                                             \
    }
LSM_PACK_LAUNCHER(16);
LSM_PACK_LAUNCHER(32);
#if LSM_IN_PART(7)
LSM_PACK_LAUNCHER_BODY(16)
#endif
#if LSM_IN_PART(8)
LSM_PACK_LAUNCHER_BODY(32)
#endif

#if LSM_IN_PART(6)
int team_max_cells_per_wave(const DevMech &m) {
    if (!m.img.team_ok) return 1;
    return m.ne <= 16 ? 4 : (m.ne <= 32 ? 2 : 1);
}

int team_waves_per_cu_cpw(const DevMech &m, int cpw) {
    return cpw > 1 ? 4 * PACK_WAVES_PER_SIMD : team_waves_per_cu(m);
}

bool team_shape_fits_cpw(const DevMech &m, int cpw, int wpb, int blocks_per_cu, size_t lds_cu) {
    if (cpw == 1) return team_shape_fits(m, wpb, blocks_per_cu, lds_cu);
    if ((cpw != 2 && cpw != 4) || cpw > team_max_cells_per_wave(m) || wpb < 1 || wpb > 4 || blocks_per_cu < 1)
        return false;
    const size_t wave = pack_wave_bytes(64 / cpw, m.ne);
    return (size_t)blocks_per_cu * ((size_t)m.img.bytes + (size_t)wpb * wave) <= lds_cu;
}

int launch_team_cpw(const DevMech &m, const KParams &k, int64_t n, const double *dt, const double *yin,
                    double *yout, int *status, double *q, int *counters, unsigned *queue, const int *order, int grid,
                    int wpb, int cpw, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    if (cpw == 1) return launch_team(m, k, n, dt, yin, yout, status, q, counters, queue, order, grid, wpb, stream);
    if (wpb < 1 || wpb > 4 || cpw > team_max_cells_per_wave(m)) return -2;
    if (cpw == 4) return launch_pack_16(m, k, n, dt, yin, yout, status, q, counters, queue, order, grid, wpb, s);
    if (cpw == 2) return launch_pack_32(m, k, n, dt, yin, yout, status, q, counters, queue, order, grid, wpb, s);
    return -2;
}

int launch_team(const DevMech &m, const KParams &k, int64_t n, const double *dt, const double *yin, double *yout,
                int *status, double *q, int *counters, unsigned *queue, const int *order, int grid, int wpb,
                void *stream) {
    hipStream_t s = (hipStream_t)stream;
    if (wpb < 1 || wpb > 4) return -2;
    if (m.ne <= 16) return launch_team_16(m, k, n, dt, yin, yout, status, q, counters, queue, order, grid, wpb, s);
    if (m.ne <= 32) return launch_team_32(m, k, n, dt, yin, yout, status, q, counters, queue, order, grid, wpb, s);
    if (m.ne <= 48) return launch_team_48(m, k, n, dt, yin, yout, status, q, counters, queue, order, grid, wpb, s);
    if (m.ne <= 56) return launch_team_56(m, k, n, dt, yin, yout, status, q, counters, queue, order, grid, wpb, s);
    if (m.ne <= 64) return launch_team_64(m, k, n, dt, yin, yout, status, q, counters, queue, order, grid, wpb, s);
    return -2;
}

// Thermodynamic part of the per-cell property refresh (solvers/laminarSMOKE/properties.H:41-131):
// one thread per cell, species rows streamed from HBM (two passes: MW, then the mole
// fractions and the mixture sums).  out rows: MW [kg/kmol], cTot = p/(RT) [kmol/m3],
// psi = cTot MW / p [s2/m2], h [J/kg], cp [J/kg/K], cv = (cp_molar - R)/MW [J/kg/K];
// x (may be null): mole fractions; cpk (may be null): species cp_k / MW_k [J/kg/K].
__global__ void __launch_bounds__(256) properties_kernel(DevMech M, double p_default, int64_t n,
                                                         const double *__restrict__ y, const double *__restrict__ p,
                                                         double *__restrict__ out, double *__restrict__ x,
                                                         double *__restrict__ cpk) {
    const double *sp = M.sp;
    const int P = M.ns_pad, ns = M.ns;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const double T = y[(int64_t)ns * n + c];
        const double pc = p ? p[c] : p_default;
        double u = 0.;
#pragma unroll 4
        for (int k = 0; k < ns; k++) u += y[(int64_t)k * n + c] / sp[SP_MW * P + k];
        const double MW = 1. / u;
        const double T2 = T * T, T3 = T2 * T, T4 = T3 * T, uT = 1. / T;
        double cpx = 0., hx = 0.;
#pragma unroll 4
        for (int k = 0; k < ns; k++) {
            const double mwk = sp[SP_MW * P + k];
            const double xk = (MW * y[(int64_t)k * n + c]) / mwk;
            const bool hi = T > sp[SP_TMID * P + k];
            const int rc = hi ? SP_CP_HT : SP_CP_LT, rh = hi ? SP_DH_HT : SP_DH_LT;
            const double cp = sp[rc * P + k] + T * sp[(rc + 1) * P + k] + T2 * sp[(rc + 2) * P + k] +
                              T3 * sp[(rc + 3) * P + k] + T4 * sp[(rc + 4) * P + k];
            const double h = sp[rh * P + k] + T * sp[(rh + 1) * P + k] + T2 * sp[(rh + 2) * P + k] +
                             T3 * sp[(rh + 3) * P + k] + T4 * sp[(rh + 4) * P + k] + uT * sp[(rh + 5) * P + k];
            cpx += xk * cp;
            hx += xk * h;
            if (x) x[(int64_t)k * n + c] = xk;
            if (cpk) cpk[(int64_t)k * n + c] = (cp * RGAS) / mwk;
        }
        const double cTot = pc / RGAS / T;
        out[c] = MW;
        out[n + c] = cTot;
        out[2 * n + c] = cTot * MW / pc;
        out[3 * n + c] = (hx * RGAS * T) / MW;
        out[4 * n + c] = (cpx * RGAS) / MW;
        out[5 * n + c] = (cpx * RGAS - RGAS) / MW;
    }
}

// Transport part of the property refresh (solvers/laminarSMOKE/properties.H:101-192):
// DynamicViscosity, ThermalConductivity, MassDiffusionCoefficients and
// ThermalDiffusionRatios of TransportPropertiesMap_CHEMKIN (maps/TransportPropertiesMap_CHEMKIN.hpp:
// 690-1117) per cell, with the solver's options: simplified mu and lambda, Lewis-number
// diffusivities, soot-diffusivity corrections and Soret coefficients.  One wave per cell;
// lane k handles species k, k + 64, ...; the wave's mole fractions and 1/sqrt(eta) sit in
// LDS.  Each species sums over its partners in increasing index order, the order in which
// the reference's pair loops add to it; the sums over species are wave reductions.  The
// Wilke factor of an earlier partner is formed as sqrt(eta_k/eta_q) / r14 rather than the
// reference's 1/(sqrt(eta_q/eta_k) r14): equal up to rounding, and without a division.  With share_pairs, each binary 1/D_jk is
// evaluated once (pairs dealt round-robin to the lanes) into a per-wave LDS table that
// both partners read; otherwise each partner evaluates it (the same value either way).
__device__ __forceinline__ int pair_index(int a, int b, int ns) {   // a < b
    return a * ns - a * (a + 1) / 2 + (b - a - 1);
}

__global__ void __launch_bounds__(256) transport_kernel(DevMech M, DevTransport Tr, double p_default, int64_t n,
                                                        const double *__restrict__ y, const double *__restrict__ p,
                                                        double *__restrict__ mu, double *__restrict__ lam,
                                                        double *__restrict__ dmix, double *__restrict__ teta,
                                                        int share_pairs) {
    extern __shared__ double tsh[];
    const int ns = M.ns, P = M.ns_pad, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int np = ns * (ns - 1) / 2;
    const size_t per_wave = (size_t)3 * ns + (share_pairs ? (size_t)np : 0);
    double *sx = tsh + (size_t)wv * per_wave, *sq = sx + ns, *s3 = sq + ns;   // x, 1/sqrt(eta) then D, eta / corrected x
    double *sp = s3 + ns;                                                     // 1/D of each pair (share_pairs)
    const bool lewis = Tr.dmodel == LSM_DIFFUSIVITY_LEWIS_NUMBERS;
    const bool need_cp = lewis || Tr.simplified;               // cpCells (properties.H:109-113)
    const bool need_d = dmix || (teta && Tr.soret);
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t c = (int64_t)blockIdx.x * (blockDim.x >> 6) + wv; c < n; c += nw) {
        const double T = y[(int64_t)ns * n + c];
        const double pc = p ? p[c] : p_default;
        const double logT = log(T), logT2 = logT * logT, logT3 = logT * logT2;
        // mole fractions (ThermodynamicsMap_CHEMKIN::MoleFractions_From_MassFractions)
        double u = 0.;
        for (int k = lane; k < ns; k += 64) u += y[(int64_t)k * n + c] / M.sp[SP_MW * P + k];
        const double MW = 1. / wave_sum(u);
        double l1 = 0., l2 = 0., cpx = 0.;
        for (int k = lane; k < ns; k += 64) {
            const double xk = (MW * y[(int64_t)k * n + c]) / M.sp[SP_MW * P + k];
            const double *a = Tr.lam + 4 * k, *e = Tr.eta + 4 * k;
            double sl = a[0];
            sl += a[1] * logT; sl += a[2] * logT2; sl += a[3] * logT3;
            double se = e[0];
            se += e[1] * logT; se += e[2] * logT2; se += e[3] * logT3;
            const double lk = exp(sl), ek = exp(se);
            sx[k] = xk;
            sq[k] = 1. / sqrt(ek);
            s3[k] = ek;
            l1 += xk * lk;
            l2 += xk / lk;
            if (need_cp) {                                     // cp_k / R (ThermodynamicsMap_CHEMKIN::cp_over_R)
                const bool hi = T > M.sp[SP_TMID * P + k];
                const int rc = hi ? SP_CP_HT : SP_CP_LT;
                const double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
                cpx += xk * (M.sp[rc * P + k] + T * M.sp[(rc + 1) * P + k] + T2 * M.sp[(rc + 2) * P + k] +
                             T3 * M.sp[(rc + 3) * P + k] + T4 * M.sp[(rc + 4) * P + k]);
            }
        }
        wave_sync();
        const double L1 = wave_sum(l1), L2 = wave_sum(l2);
        const double cp = need_cp ? (wave_sum(cpx) * RGAS) / MW : 0.;    // J/kg/K
        double lam_c = 0.50 * (L1 + 1. / L2);                          // lambdaMix (:947-957)
        double mu_c;
        if (Tr.simplified) {                                           // properties.H:130-139
            mu_c = 1.8405e-5 * pow(T / 300., 0.6759);
            lam_c = mu_c * cp / 0.7;
        } else {
            // etaMix (:959-1030)
            double m1 = 0., m2 = 0.;
            for (int k = lane; k < ns; k += 64) {
                const double xk = sx[k], ek = s3[k];
                if (Tr.model == LSM_VISCOSITY_WILKE) {
                    // partner-major matrices: lane k reads [q ns + k], contiguous across the
                    // wave; the diagonal weight is 0, so q = k adds nothing and needs no branch
                    const double sqk = sqrt(ek);
                    const double *wwk = Tr.ww + k, *wrk = Tr.wr + k;
                    double sumK = xk;
                    for (int q = 0; q < ns; q++) {
                        const double e = sqk * sq[q] * wrk[(size_t)q * ns];
                        sumK += sx[q] * wwk[(size_t)q * ns] * (1. + e) * (1. + e);
                    }
                    m1 += xk * ek / sumK;
                } else if (Tr.model == LSM_VISCOSITY_HERNING) {
                    const double sm = sqrt(Tr.mw[k]);
                    m1 += xk * ek * sm;
                    m2 += xk * sm;
                } else {
                    m1 += xk * ek;
                    m2 += xk / ek;
                }
            }
            const double M1 = wave_sum(m1), M2 = wave_sum(m2);
            mu_c = Tr.model == LSM_VISCOSITY_WILKE ? M1
                                                   : (Tr.model == LSM_VISCOSITY_HERNING ? M1 / M2 : 0.50 * (M1 + 1. / M2));
        }
        if (lane == 0) {
            mu[c] = mu_c;
            lam[c] = lam_c;
        }
        // Diffusion coefficients into sq (1/sqrt(eta) is no longer read): gamma() and gammaMix
        // (:798-859, 1032-1059), or lambda / (p psi) / cp / Le_k (properties.H:153-164)
        if (need_d) {
            wave_sync();
            if (lewis) {
                const double cTot = pc / RGAS / T, psi = cTot * MW / pc;
                const double coef = lam_c / (pc * psi) / cp;
                for (int k = lane; k < ns; k += 64) sq[k] = coef / Tr.lewis[k];
            } else {
                const double thr = 1.e-14, sum_thr = 1. + thr * (double)ns, P_bar = pc / 100000.;
                double mc = 0.;
                for (int k = lane; k < ns; k += 64) {
                    const double xc = (sx[k] + thr) / sum_thr;
                    s3[k] = xc;
                    mc += xc * Tr.mw[k];
                }
                if (share_pairs)
                    for (int q = lane; q < np; q += 64) {
                        const double *f = Tr.gam + 4 * q;
                        double sum = f[0];
                        sum += f[1] * logT; sum += f[2] * logT2; sum += f[3] * logT3;
                        sp[q] = P_bar / exp(sum);
                    }
                wave_sync();
                const double MWc = wave_sum(mc);
                for (int k = lane; k < ns; k += 64) {
                    double sdc = 0.;
                    for (int q = 0; q < ns; q++) {
                        if (q == k) continue;
                        const int i = pair_index(q < k ? q : k, q < k ? k : q, ns);
                        double dinv;
                        if (share_pairs) {
                            dinv = sp[i];
                        } else {
                            const double *f = Tr.gam + 4 * i;
                            double sum = f[0];
                            sum += f[1] * logT; sum += f[2] * logT2; sum += f[3] * logT3;
                            dinv = P_bar / exp(sum);
                        }
                        sdc += s3[q] * dinv;      // x_q / D_kq
                    }
                    sq[k] = (MWc - s3[k] * Tr.mw[k]) / (MWc * sdc);
                }
            }
            wave_sync();
        }
        // soot corrections (properties.H:166-177), then the outputs; tetaMix (:768-796,
        // 1103-1117): theta_kj x_k x_j for the light species, or Dsoret (properties.H:180-192)
        for (int k = lane; k < ns; k += 64) {
            double D = 0.;
            if (need_d) {
                D = sq[k];
                if (Tr.soot_ref >= 0) {
                    const int si = Tr.soot_slot[k];
                    if (si >= 0) D = sq[Tr.soot_ref] * Tr.soot_corr[si];
                }
                if (dmix) dmix[(int64_t)k * n + c] = D;
            }
            if (teta) {
                const int sl = Tr.teta_slot[k];
                double tm = 0.;
                if (sl >= 0) {
                    const double T2 = T * T, T3 = T * T2, xk = sx[k];
                    const double *f = Tr.teta + (size_t)4 * sl * ns;
                    for (int j = 0; j < ns; j++) {
                        double sum = f[4 * j];
                        sum += f[4 * j + 1] * T; sum += f[4 * j + 2] * T2; sum += f[4 * j + 3] * T3;
                        tm += sum * xk * sx[j];
                    }
                    if (Tr.soret) tm = D * tm * M.sp[SP_MW * P + k] / MW;
                }
                teta[(int64_t)k * n + c] = tm;
            }
        }
        wave_sync();
    }
}

int launch_transport(const DevMech &m, const DevTransport &t, double p_default, int64_t n, const double *y,
                     const double *p, double *mu, double *lam, double *dmix, double *teta, int num_cu, void *stream) {
    if (n <= 0) return 0;
    const size_t base = (size_t)3 * m.ns, np = (size_t)m.ns * (m.ns - 1) / 2;
    const int share = (base + np) * sizeof(double) * 4 <= 65536 ? 1 : 0;
    const size_t lds = (base + (share ? np : 0)) * sizeof(double) * 4;
    if (lds > 65536) return -2;
    const int64_t blocks = (n + 3) / 4;
    const int g = (int)(blocks < (int64_t)num_cu * 16 ? blocks : (int64_t)num_cu * 16);
    hipLaunchKernelGGL(transport_kernel, dim3(g), dim3(256), lds, (hipStream_t)stream, m, t, p_default, n, y, p, mu,
                       lam, dmix, teta, share);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_properties(const DevMech &m, double p_default, int64_t n, const double *y, const double *p, double *out,
                      double *x, double *cpk, int num_cu, void *stream) {
    if (n <= 0) return 0;
    const int64_t blocks = (n + 255) / 256;
    const int g = (int)(blocks < (int64_t)num_cu * 64 ? blocks : (int64_t)num_cu * 64);
    hipLaunchKernelGGL(properties_kernel, dim3(g), dim3(256), 0, (hipStream_t)stream, m, p_default, n, y, p, out, x,
                       cpk);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_equations(const DevMech &m, const KParams &k, int64_t n, const double *yin, double *f, double *J,
                     int grid, int wpb, void *stream) {
    if (wpb < 1 || wpb > 4 || m.ne > 64) return -2;
    const int64_t need = (n + wpb - 1) / wpb;
    const int g = (int)(need < grid ? need : grid);
    if (g < 1) return 0;
    const size_t lds = (size_t)m.img.bytes + (size_t)wpb * team_wave_bytes(m.ne);
    const void *fn = m.xk ? (const void *)equations_kernel<true> : (const void *)equations_kernel<false>;
    if (lds > 65536 && hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return -1;
    if (m.xk)
        hipLaunchKernelGGL(equations_kernel<true>, dim3(g), dim3(64 * wpb), lds, (hipStream_t)stream, m, k, n, yin, f,
                           J);
    else
        hipLaunchKernelGGL(equations_kernel<false>, dim3(g), dim3(64 * wpb), lds, (hipStream_t)stream, m, k, n, yin, f,
                           J);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

size_t wide_lds_bytes(const DevMech &m) { return WideCtx<false>::lds_bytes(m.ne); }
size_t wide_jac_doubles(const DevMech &m) { return (size_t)2 * m.ne * m.ne; }   // J, then the FP64 LU

int launch_wide(const DevMech &m, const KParams &k, int64_t n, const double *dt, const double *yin, double *yout,
                int *status, double *q, int *counters, unsigned *queue, const int *order, int grid, double *jws,
                void *stream) {
    if (!m.img.wide_ok || m.ne > WIDE_MAX_NEQ) return -2;
    const size_t lds = wide_lds_bytes(m);
    static size_t lds_set[2] = {65536, 65536};
    const void *fn = m.xk ? (const void *)wide_kernel<true> : (const void *)wide_kernel<false>;
    if (lds > lds_set[m.xk]) {
        if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) return -1;
        lds_set[m.xk] = lds;
    }
    if (m.xk)
        hipLaunchKernelGGL(wide_kernel<true>, dim3(grid), dim3(WIDE_THREADS), lds, (hipStream_t)stream, m, k, n, dt,
                           yin, yout, status, q, counters, queue, order, jws);
    else
        hipLaunchKernelGGL(wide_kernel<false>, dim3(grid), dim3(WIDE_THREADS), lds, (hipStream_t)stream, m, k, n, dt,
                           yin, yout, status, q, counters, queue, order, jws);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

size_t plain_work_doubles(const DevMech &m, int64_t slots) {
    const int64_t ne = m.ne;
    const int64_t per = (int64_t)NVEC * ne + 2 * ne * ne + 3 * (int64_t)m.ns + m.nr + ne;  // piv: 2 ints = 1 double
    return (size_t)(per * slots);
}

int launch_plain(const DevMech &m, const KParams &k, int64_t n, const double *dt, const double *yin, double *yout,
                 int *status, double *q, int *counters, double *work, int64_t slots, void *stream) {
    const int bs = 256;
    const int grid = (int)((slots + bs - 1) / bs);
    if (m.xk)
        hipLaunchKernelGGL(plain_kernel<true>, dim3(grid), dim3(bs), 0, (hipStream_t)stream, m, k, n, dt, yin, yout,
                           status, q, counters, work, slots);
    else
        hipLaunchKernelGGL(plain_kernel<false>, dim3(grid), dim3(bs), 0, (hipStream_t)stream, m, k, n, dt, yin, yout,
                           status, q, counters, work, slots);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Sort keys for a cost order: key = BDF steps of the cell in a previous call.
__global__ void order_keys_kernel(const int *counters, int64_t n, unsigned *keys, int *vals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        keys[i] = counters ? (unsigned)counters[i * LSM_NCOUNTERS_DEV] : 0u;
        vals[i] = (int)i;
    }
}

// Strided sample of a batch (cells i * stride), compacted: ys [ne][m], dts [m].
__global__ void gather_cells_kernel(const double *y, const double *dt, int64_t n, int ne, int64_t m, int64_t stride,
                                    double *ys, double *dts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) {
        const int64_t src = i * stride;
        for (int e = 0; e < ne; e++) ys[(int64_t)e * m + i] = y[(int64_t)e * n + src];
        dts[i] = dt[src];
    }
}

int launch_order_keys(const int *counters, int64_t n, unsigned *keys, int *vals, void *stream) {
    const int bs = 256;
    hipLaunchKernelGGL(order_keys_kernel, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, (hipStream_t)stream,
                       counters, n, keys, vals);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_gather_cells(const double *y, const double *dt, int64_t n, int ne, int64_t m, int64_t stride, double *ys,
                        double *dts, void *stream) {
    const int bs = 256;
    hipLaunchKernelGGL(gather_cells_kernel, dim3((unsigned)((m + bs - 1) / bs)), dim3(bs), 0, (hipStream_t)stream, y,
                       dt, n, ne, m, stride, ys, dts);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ void fill_kernel(double *p, int64_t n, double v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int launch_fill(double *p, int64_t n, double v, void *stream) {
    const int bs = 256;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, (hipStream_t)stream, p, n, v);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_reset_queue(unsigned *queue, void *stream) {
    hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(1), 0, (hipStream_t)stream, queue);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
#endif  // LSM_IN_PART(6)

}  // namespace lsm
