"""Independent dense NumPy formulation of the constant-pressure reactor, used to pin
the C oracle.  It is written from the textbook definitions, not from the oracle's
loops:

* NASA-7: cp/R as a polynomial; h/R and s/R from its antiderivatives
  (numpy.polynomial), plus the integration constants a6, a7;
* Kc from the standard Gibbs energy: Kc = exp(-sum_k nu_k g_k/RT) (Patm/RT)^dn;
* rates r_j = k_j (prod c^nu' - prod c^nu'' / Kc) with dense order matrices,
  third-body M_j = sum_k eff_jk c_k, Lindemann/Troe fall-off blending;
* R = nu r; dY/dt = MW R / rho; dT/dt = -sum_k h_k R_k / (rho cp).
"""
import numpy as np
from numpy.polynomial import polynomial as P

R = 8314.4621
PATM = 101325.0


def nasa_props(mech, T):
    a = mech.nasa
    hi = T > a[:, 16]
    c = np.where(hi[:, None], a[:, 0:7], a[:, 7:14])          # a1..a7 of the active range
    cp = np.array([P.polyval(T, c[k, :5]) for k in range(mech.ns)])
    h_R = np.array([P.polyval(T, P.polyint(c[k, :5])) for k in range(mech.ns)]) + c[:, 5]
    s_R = c[:, 0] * np.log(T) + np.array([P.polyval(T, P.polyint(c[k, 1:5])) for k in range(mech.ns)]) + c[:, 6]
    return cp, h_R / T, s_R


def _dense(mech):
    ns, nr = mech.ns, mech.nr
    of, ob = np.zeros((nr, ns)), np.zeros((nr, ns))
    for j in range(nr):
        for p in range(mech.reac_ptr[j], mech.reac_ptr[j + 1]):
            of[j, mech.reac_sp[p]] = mech.reac_nu[p]
        for p in range(mech.prod_ptr[j], mech.prod_ptr[j + 1]):
            ob[j, mech.prod_sp[p]] = mech.prod_nu[p]
    return of, ob


def rates(mech, T, c):
    """Net rates r[nr] and formation rates R[ns] at (T, c)."""
    of, ob = _dense(mech)
    nu = (ob - of).T
    cp, h_RT, s_R = nasa_props(mech, T)
    g_RT = h_RT - s_R
    A = np.exp(mech.lnA) * np.where(mech.rflags & 2, -1.0, 1.0)
    k = A * T ** mech.beta * np.exp(-mech.E_over_R / T)
    dn = (ob - of).sum(axis=1)
    Kc = np.exp(-(nu.T @ g_RT)) * (PATM / (R * T)) ** dn
    for s, j in enumerate(mech.tb_rxn):
        eff = np.ones(mech.ns)
        for p in range(mech.tb_ptr[s], mech.tb_ptr[s + 1]):
            eff[mech.tb_sp[p]] = mech.tb_eff_m1[p] + 1.0
        k[j] *= eff @ c
    for s, j in enumerate(mech.fo_rxn):
        if mech.fo_single[s] >= 0:
            M = c[mech.fo_single[s]] + 1e-16
        else:
            eff = np.ones(mech.ns)
            for p in range(mech.fo_ptr[s], mech.fo_ptr[s + 1]):
                eff[mech.fo_sp[p]] = mech.fo_eff_m1[p] + 1.0
            M = eff @ c
        lA, b, e = mech.fo_inf[s]
        kinf = np.exp(lA) * T ** b * np.exp(-e / T)
        k0 = k[j]
        Pr = k0 * M / kinf
        F = 1.0
        if mech.fo_type[s] == 1:
            a, T3, T1, T2 = mech.fo_troe[s]
            Fc = (1 - a) * np.exp(-T / T3) + a * np.exp(-T / T1) + (np.exp(-T2 / T) if T2 != 0 else 0.0)
            F = troe_F(Fc, Pr)
        k[j] = kinf * Pr / (1 + Pr) * F
    fwd = np.prod(c[None, :] ** of, axis=1)
    rev = np.prod(c[None, :] ** ob, axis=1)
    rev_on = (mech.rflags & 1).astype(bool)
    r = k * (fwd - np.where(rev_on, rev / Kc, 0.0))
    return r, nu @ r


def troe_F(Fc, Pr):
    """Gilbert-Luther-Troe broadening factor."""
    lFc = np.log10(Fc)
    if Pr <= 1e-32:
        return 10.0 ** (lFc / (1.0 + (1.0 / 0.14) ** 2))
    c = -0.4 - 0.67 * lFc
    n = 0.75 - 1.27 * lFc
    x = (np.log10(Pr) + c) / (n - 0.14 * (np.log10(Pr) + c))
    return 10.0 ** (lFc / (1.0 + x * x))


def rhs(mech, y, P=PATM, energy=True):
    ns = mech.ns
    Y, T = y[:ns], y[ns]
    mw = mech.mw
    W = 1.0 / np.sum(Y / mw)
    ctot = P / (R * T)
    c = ctot * W * Y / mw
    rho = ctot * W
    r, Rf = rates(mech, T, c)
    dY = mw * Rf / rho
    cp, h_RT, _ = nasa_props(mech, T)
    x = W * Y / mw
    cp_mass = (x @ cp) * R / W
    QR = -(Rf @ h_RT) * R * T
    dT = QR / (rho * cp_mass) if energy else 0.0
    return np.concatenate([dY, [dT]]), QR
