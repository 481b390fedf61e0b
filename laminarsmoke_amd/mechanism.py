"""Mechanism tables and the seeded synthetic mechanisms used by the workloads.

A :class:`Mechanism` holds the flat arrays described in ``include/lsm_mechanism.h``
(SI units: kmol, m3, s, J, K).  It is what OpenSMOKE++'s
``ThermodynamicsMap_CHEMKIN`` / ``KineticsMap_CHEMKIN`` / ``StoichiometricMap``
hold after reading ``kinetics.xml`` (maps/ThermodynamicsMap_CHEMKIN.hpp:263-343,
maps/KineticsMap_CHEMKIN.hpp:311-1165, maps/StoichiometricMap.hpp:78-254).

:func:`synthetic_mechanism` builds element-balanced C/H/O/N mechanisms:

* species are C_c H_h O_o formulas plus inert N2, molecular weights from atomic
  weights;
* NASA-7 coefficients are least-squares fits of a smooth cp(T) model over the
  300/1000/3500 K ranges, with cp continuous at T_mid and the a6/a7 constants
  fixed by h(298.15 K), s(298.15 K) and continuity of h and s at T_mid;
* reactions are reversible; A is log-uniform in [1e6, 1e14] (cgs, converted to
  SI with (1e-3)^(order-1) like kernel/kinetics/ReactionPolicy_CHEMKIN.hpp:1378-1400),
  beta uniform in [-2, 2], Ea uniform in [0, 40] kcal/mol, converted with the
  reference's J_from_cal = 4.18443 (math/Conversions.h:112) and R = 8.3144621
  J/mol/K (math/PhysicalConstants.h:20);
* third-body efficiencies are uniform in [0.5, 12] and stored as eff - 1
  (maps/KineticsMap_CHEMKIN.hpp:655-661); fall-off reactions carry LOW (direct)
  and HIGH (k_inf) Arrhenius sets and Troe parameters.
"""
from __future__ import annotations

import ctypes
import itertools
from dataclasses import dataclass, field

import numpy as np

from ._abi import LsmMechanism

ELEMENTS = ("C", "H", "O", "N")
ATOMIC_WEIGHTS = np.array([12.011, 1.008, 15.999, 14.007])
R_J_KMOL = 8314.4621
R_J_MOL = 8.3144621
J_FROM_CAL = 4.18443
RXN_REVERSIBLE = 1
RXN_NEGATIVE_A = 2
RXN_EXPLICIT_REV = 4      # with RXN_REVERSIBLE: reverse k from rev_arr (REV)
FALLOFF_LINDEMANN = 0
FALLOFF_TROE = 1
FALLOFF_SRI = 2
FALLOFF_CHEMACT = 8       # bit ORed onto the form: chemically activated (HIGH, CABR)


@dataclass
class Mechanism:
    """Flat mechanism tables (0-based, SI).  Field meaning: include/lsm_mechanism.h."""

    species: list
    elements: np.ndarray            # [ns, 4] atoms of C, H, O, N
    nasa: np.ndarray                # [ns, 18]
    lnA: np.ndarray                 # [nr]
    beta: np.ndarray                # [nr]
    E_over_R: np.ndarray            # [nr]
    rflags: np.ndarray              # [nr] int32
    reac_ptr: np.ndarray
    reac_sp: np.ndarray
    reac_nu: np.ndarray
    prod_ptr: np.ndarray
    prod_sp: np.ndarray
    prod_nu: np.ndarray
    tb_rxn: np.ndarray
    tb_ptr: np.ndarray
    tb_sp: np.ndarray
    tb_eff_m1: np.ndarray
    fo_rxn: np.ndarray
    fo_type: np.ndarray
    fo_inf: np.ndarray              # [nfo, 3]
    fo_troe: np.ndarray             # [nfo, 4]
    fo_single: np.ndarray
    fo_ptr: np.ndarray
    fo_sp: np.ndarray
    fo_eff_m1: np.ndarray
    name: str = "mechanism"
    fo_sri: object = None             # [nfo, 5] SRI a, b, c, d, e (rows of FALLOFF_SRI reactions)
    ord_ptr: object = None            # [nr+1] explicit forward orders (FORD), CSR; empty row: orders = coefficients
    ord_sp: object = None
    ord_val: object = None
    rev_arr: object = None            # [nr, 3] explicit reverse lnA, beta, E_over_R (RXN_EXPLICIT_REV rows)
    pl_rxn: object = None             # [npl] PLOG reactions
    pl_ptr: object = None             # [npl+1] their pressure levels
    pl_par: object = None             # [levels, 4] P [Pa], lnA, beta, E_over_R
    ch_rxn: object = None             # [nch] Chebyshev reactions
    ch_dims: object = None            # [nch, 2] N (temperature), M (pressure)
    ch_lim: object = None             # [nch, 5] Tmin, Tmax [K], Pmin, Pmax [Pa], conv
    ch_ptr: object = None             # [nch+1] first coefficient of each
    ch_a: object = None               # coefficients a_nm, row-major
    element_names: tuple = ELEMENTS    # columns of `elements`
    atomic_weights: object = None      # [n_elements] kg/kmol (default: C, H, O, N)
    _struct: object = field(default=None, repr=False)

    @property
    def ns(self) -> int:
        return len(self.species)

    @property
    def nr(self) -> int:
        return int(self.lnA.shape[0])

    @property
    def neq(self) -> int:
        return self.ns + 1

    @property
    def mw(self) -> np.ndarray:
        return self.nasa[:, 17].copy()

    def __post_init__(self):
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        for name in ("nasa", "lnA", "beta", "E_over_R", "reac_nu", "prod_nu",
                     "tb_eff_m1", "fo_inf", "fo_troe", "fo_eff_m1"):
            setattr(self, name, f64(getattr(self, name)))
        for name in ("rflags", "reac_ptr", "reac_sp", "prod_ptr", "prod_sp", "tb_rxn",
                     "tb_ptr", "tb_sp", "fo_rxn", "fo_type", "fo_single", "fo_ptr", "fo_sp"):
            setattr(self, name, i32(getattr(self, name)))
        self.elements = np.ascontiguousarray(self.elements, dtype=np.float64)
        if self.fo_sri is None:
            self.fo_sri = np.tile([0.0, 0.0, 1.0, 1.0, 0.0], (len(self.fo_rxn), 1))
        self.fo_sri = np.ascontiguousarray(np.asarray(self.fo_sri, dtype=np.float64).reshape(len(self.fo_rxn), 5))
        if self.ord_ptr is None:
            self.ord_ptr, self.ord_sp, self.ord_val = np.zeros(self.nr + 1), [], []
        self.ord_ptr = np.ascontiguousarray(self.ord_ptr, dtype=np.int32)
        self.ord_sp = np.ascontiguousarray(self.ord_sp, dtype=np.int32)
        self.ord_val = np.ascontiguousarray(self.ord_val, dtype=np.float64)
        if self.rev_arr is None:
            self.rev_arr = np.zeros((self.nr, 3))
        self.rev_arr = np.ascontiguousarray(np.asarray(self.rev_arr, dtype=np.float64).reshape(self.nr, 3))
        if self.pl_rxn is None:
            self.pl_rxn, self.pl_ptr, self.pl_par = [], [0], np.zeros((0, 4))
        self.pl_rxn = np.ascontiguousarray(self.pl_rxn, dtype=np.int32)
        self.pl_ptr = np.ascontiguousarray(self.pl_ptr, dtype=np.int32)
        self.pl_par = np.ascontiguousarray(np.asarray(self.pl_par, dtype=np.float64).reshape(-1, 4))
        if self.ch_rxn is None:
            self.ch_rxn, self.ch_dims, self.ch_lim, self.ch_ptr, self.ch_a = [], [], [], [0], []
        self.ch_rxn = np.ascontiguousarray(self.ch_rxn, dtype=np.int32)
        self.ch_dims = np.ascontiguousarray(np.asarray(self.ch_dims, dtype=np.int32).reshape(-1, 2))
        self.ch_lim = np.ascontiguousarray(np.asarray(self.ch_lim, dtype=np.float64).reshape(-1, 5))
        self.ch_ptr = np.ascontiguousarray(self.ch_ptr, dtype=np.int32)
        self.ch_a = np.ascontiguousarray(self.ch_a, dtype=np.float64)
        if self.atomic_weights is None:
            self.atomic_weights = ATOMIC_WEIGHTS
        self.atomic_weights = np.asarray(self.atomic_weights, dtype=np.float64)
        self.validate()

    def validate(self):
        ns, nr = self.ns, self.nr
        assert self.nasa.shape == (ns, 18)
        for side in ("reac", "prod"):
            ptr = getattr(self, side + "_ptr")
            sp = getattr(self, side + "_sp")
            assert ptr.shape == (nr + 1,) and ptr[0] == 0 and ptr[-1] == sp.shape[0]
            assert np.all(np.diff(ptr) >= 1)
            assert np.all((sp >= 0) & (sp < ns))
            for j in range(nr):
                s = sp[ptr[j]:ptr[j + 1]]
                assert len(set(s.tolist())) == len(s), "species repeated on one side"
        assert self.tb_ptr.shape == (len(self.tb_rxn) + 1,)
        assert self.fo_ptr.shape == (len(self.fo_rxn) + 1,)
        assert self.fo_inf.shape == (len(self.fo_rxn), 3)
        assert self.fo_troe.shape == (len(self.fo_rxn), 4)
        assert self.pl_ptr.shape == (len(self.pl_rxn) + 1,) and self.pl_ptr[-1] == self.pl_par.shape[0]
        nch = len(self.ch_rxn)
        assert self.ch_dims.shape == (nch, 2) and self.ch_lim.shape == (nch, 5)
        assert self.ch_ptr.shape == (nch + 1,) and self.ch_ptr[-1] == self.ch_a.shape[0]

    def struct(self) -> LsmMechanism:
        """ctypes view of the tables (valid while this object lives)."""
        if self._struct is None:
            P = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))
            d, i = ctypes.c_double, ctypes.c_int32
            s = LsmMechanism()
            s.ns, s.nr = self.ns, self.nr
            s.nasa = P(self.nasa, d)
            s.lnA, s.beta, s.E_over_R = P(self.lnA, d), P(self.beta, d), P(self.E_over_R, d)
            s.rflags = P(self.rflags, i)
            s.reac_ptr, s.reac_sp, s.reac_nu = P(self.reac_ptr, i), P(self.reac_sp, i), P(self.reac_nu, d)
            s.prod_ptr, s.prod_sp, s.prod_nu = P(self.prod_ptr, i), P(self.prod_sp, i), P(self.prod_nu, d)
            s.ntb = len(self.tb_rxn)
            s.tb_rxn, s.tb_ptr, s.tb_sp = P(self.tb_rxn, i), P(self.tb_ptr, i), P(self.tb_sp, i)
            s.tb_eff_m1 = P(self.tb_eff_m1, d)
            s.nfo = len(self.fo_rxn)
            s.fo_rxn, s.fo_type = P(self.fo_rxn, i), P(self.fo_type, i)
            s.fo_inf, s.fo_troe = P(self.fo_inf, d), P(self.fo_troe, d)
            s.fo_single, s.fo_ptr, s.fo_sp = P(self.fo_single, i), P(self.fo_ptr, i), P(self.fo_sp, i)
            s.fo_eff_m1 = P(self.fo_eff_m1, d)
            s.fo_sri = P(self.fo_sri, d)
            s.ord_ptr, s.ord_sp, s.ord_val = P(self.ord_ptr, i), P(self.ord_sp, i), P(self.ord_val, d)
            s.rev_arr = P(self.rev_arr, d)
            s.npl = len(self.pl_rxn)
            s.pl_rxn, s.pl_ptr, s.pl_par = P(self.pl_rxn, i), P(self.pl_ptr, i), P(self.pl_par, d)
            s.nch = len(self.ch_rxn)
            s.ch_rxn, s.ch_dims, s.ch_lim = P(self.ch_rxn, i), P(self.ch_dims, i), P(self.ch_lim, d)
            s.ch_ptr, s.ch_a = P(self.ch_ptr, i), P(self.ch_a, d)
            self._struct = s
        return self._struct

    # --- dense views used by tests -------------------------------------------------
    def stoich_matrix(self) -> np.ndarray:
        """Net stoichiometric matrix nu[ns, nr] (products - reactants)."""
        nu = np.zeros((self.ns, self.nr))
        for j in range(self.nr):
            for p in range(self.reac_ptr[j], self.reac_ptr[j + 1]):
                nu[self.reac_sp[p], j] -= self.reac_nu[p]
            for p in range(self.prod_ptr[j], self.prod_ptr[j + 1]):
                nu[self.prod_sp[p], j] += self.prod_nu[p]
        return nu

    def element_mass_matrix(self) -> np.ndarray:
        """[n_elements, ns] kg of element e per kg of species k (for element-mass balances)."""
        return (self.elements * self.atomic_weights[None, :]).T / self.mw[None, :]


# ----------------------------------------------------------------------------------
# species and thermodynamics
# ----------------------------------------------------------------------------------
_CORE14 = ["H2", "O2", "H2O", "H", "O", "OH", "HO2", "H2O2", "CO", "CO2", "HCO", "CH3", "CH4", "N2"]


def _parse(formula: str) -> np.ndarray:
    out = np.zeros(4)
    i = 0
    while i < len(formula):
        el = formula[i]
        i += 1
        num = ""
        while i < len(formula) and formula[i].isdigit():
            num += formula[i]
            i += 1
        out[ELEMENTS.index(el)] += int(num) if num else 1
    return out


def _name(c, h, o) -> str:
    s = ""
    for el, n in (("C", c), ("H", h), ("O", o)):
        if n:
            s += el + (str(n) if n > 1 else "")
    return s


def species_list(ns: int) -> list:
    """Deterministic species set: the 14 core species, then C/H/O formulas by size."""
    if ns <= len(_CORE14):
        if ns != len(_CORE14):
            raise ValueError("synthetic mechanisms need ns >= 14")
        return list(_CORE14)
    names = list(_CORE14[:-1])
    seen = {tuple(_parse(n).astype(int)) for n in _CORE14}
    extra = []
    for c, o, h in itertools.product(range(0, 7), range(0, 4), range(0, 15)):
        if c + h + o == 0 or h > 2 * c + 2 or (c == 0 and o > 2):
            continue
        key = (c, h, o, 0)
        if key in seen:
            continue
        seen.add(key)
        extra.append((c + o + 0.5 * h, c, o, h, _name(c, h, o)))
    extra.sort()
    for *_, nm in extra:
        if len(names) >= ns - 1:
            break
        names.append(nm)
    names.append("N2")
    return names


def _cp_model(T, natoms, theta):
    """Smooth dimensionless cp/R(T) that rises from the low-T to the classical limit."""
    if natoms == 1:
        return 2.5 + 0.0 * T
    lo = 3.5 if natoms == 2 else 4.0
    hi = 4.5 if natoms == 2 else 3.0 * natoms - 2.0
    x = (T / theta) ** 2
    return lo + (hi - lo) * x / (1.0 + x)


def _nasa_fit(rng, formula_atoms, h298_over_R, s298_over_R):
    natoms = int(formula_atoms.sum())
    theta = rng.uniform(600.0, 1600.0)
    Tl = np.linspace(300.0, 1000.0, 60)
    Th = np.linspace(1000.0, 6000.0, 120)   # fitted past T_high so extrapolation stays smooth
    V = lambda T: np.vander(T, 5, increasing=True)
    Al, Ah = V(Tl), V(Th)
    bl, bh = _cp_model(Tl, natoms, theta), _cp_model(Th, natoms, theta)
    # least squares for [a_low(5), a_high(5)] with cp_low(1000) == cp_high(1000)
    A = np.zeros((Tl.size + Th.size, 10))
    A[:Tl.size, :5] = Al / 1000.0 ** np.arange(5)  # scaled unknowns for conditioning
    A[Tl.size:, 5:] = Ah / 1000.0 ** np.arange(5)
    b = np.concatenate([bl, bh])
    c = np.zeros(10)
    c[:5] = 1.0
    c[5:] = -1.0
    K = np.zeros((11, 11))
    K[:10, :10] = 2.0 * A.T @ A
    K[:10, 10] = c
    K[10, :10] = c
    rhs = np.concatenate([2.0 * A.T @ b, [0.0]])
    sol = np.linalg.solve(K, rhs)
    scale = 1.0 / 1000.0 ** np.arange(5)
    a_lo = sol[:5] * scale
    a_hi = sol[5:10] * scale

    def h_poly(a, T):  # h/R without a6
        return a[0] * T + a[1] * T**2 / 2 + a[2] * T**3 / 3 + a[3] * T**4 / 4 + a[4] * T**5 / 5

    def s_poly(a, T):  # s/R without a7
        return a[0] * np.log(T) + a[1] * T + a[2] * T**2 / 2 + a[3] * T**3 / 3 + a[4] * T**4 / 4

    T0, Tm = 298.15, 1000.0
    a6_lo = h298_over_R - h_poly(a_lo, T0)
    a7_lo = s298_over_R - s_poly(a_lo, T0)
    a6_hi = (h_poly(a_lo, Tm) + a6_lo) - h_poly(a_hi, Tm)
    a7_hi = (s_poly(a_lo, Tm) + a7_lo) - s_poly(a_hi, Tm)
    mw = float(formula_atoms @ ATOMIC_WEIGHTS)
    row = np.empty(18)
    row[0:5] = a_hi
    row[5] = a6_hi
    row[6] = a7_hi
    row[7:12] = a_lo
    row[12] = a6_lo
    row[13] = a7_lo
    row[14], row[15], row[16], row[17] = 300.0, 3500.0, 1000.0, mw
    return row


# ----------------------------------------------------------------------------------
# reactions
# ----------------------------------------------------------------------------------
def _loguniform(rng, lo, hi):
    return float(np.exp(rng.uniform(np.log(lo), np.log(hi))))


def synthetic_mechanism(ns: int, nr: int, n_tb: int, n_troe: int, seed: int,
                        name: str | None = None) -> Mechanism:
    """Seeded element-balanced mechanism: ns species, nr reversible reactions, of which
    n_tb are third-body and n_troe Troe fall-off reactions; the rest are elementary."""
    rng = np.random.default_rng(seed)
    names = species_list(ns)
    atoms = np.array([_parse(s) for s in names])
    key = lambda v: tuple(int(x) for x in v)

    # thermodynamics
    nasa = np.empty((ns, 18))
    h298 = np.zeros(ns)
    for k, s in enumerate(names):
        mw = float(atoms[k] @ ATOMIC_WEIGHTS)
        if s not in ("H2", "O2", "N2"):
            h298[k] = rng.uniform(-20000.0, 20000.0)       # K (i.e. +-40 kcal/mol)
        s298 = 12.0 + 3.5 * np.log(mw) + rng.uniform(0.0, 4.0)
        nasa[k] = _nasa_fit(rng, atoms[k], h298[k], s298)

    # element-vector lookup of single species and pairs (N2 kept inert)
    reactive = [k for k, s in enumerate(names) if s != "N2"]
    combos = {}
    for k in reactive:
        combos.setdefault(key(atoms[k]), []).append((k,))
    for a, b in itertools.combinations_with_replacement(reactive, 2):
        combos.setdefault(key(atoms[a] + atoms[b]), []).append((a, b))
    singles = {key(atoms[k]): k for k in reactive}

    def pick_elementary():
        while True:
            if rng.random() < 0.9:
                r = tuple(sorted(rng.choice(reactive, 2, replace=True).tolist()))
            else:
                r = (int(rng.choice(reactive)),)
            cand = [p for p in combos.get(key(atoms[list(r)].sum(0)), [])
                    if tuple(sorted(p)) != r and not (set(p) & set(r))]
            if cand:
                return r, cand[rng.integers(len(cand))]

    def pick_association():
        # A + B (+M) <=> C: pairs whose element sum is a single species
        while True:
            a, b = sorted(rng.choice(reactive, 2, replace=True).tolist())
            c = singles.get(key(atoms[a] + atoms[b]))
            if c is not None and c not in (a, b):
                return (a, b), (c,)

    def to_side(sp):
        cnt = {}
        for s in sp:
            cnt[int(s)] = cnt.get(int(s), 0) + 1
        return sorted(cnt.items())

    def arrhenius(order):
        A_cgs = _loguniform(rng, 1e6, 1e14)
        beta = rng.uniform(-2.0, 2.0)
        Ea_cal = rng.uniform(0.0, 40000.0)
        A_si = A_cgs * (1e-3) ** (order - 1)
        return np.log(A_si), beta, Ea_cal * J_FROM_CAL / R_J_MOL

    def efficiencies():
        m = int(rng.integers(3, 7))
        sp = sorted(rng.choice(ns, m, replace=False).tolist())
        return sp, [rng.uniform(0.5, 12.0) - 1.0 for _ in sp]

    kinds = ["el"] * (nr - n_tb - n_troe) + ["tb"] * n_tb + ["troe"] * n_troe
    kinds = [kinds[i] for i in rng.permutation(nr)]
    seen = set()
    lnA, beta, EoR, flags = [], [], [], []
    reac, prod = [], []
    tb_rxn, tb_sp, tb_eff = [], [], []
    fo_rxn, fo_inf, fo_troe, fo_sp, fo_eff = [], [], [], [], []
    for j, kind in enumerate(kinds):
        while True:
            if kind == "el":
                r, p = pick_elementary()
            else:
                r, p = pick_association()
            # written in the exothermic direction at 298 K, so that k_r = k_f / Kc
            # stays bounded by k_f (random Ea carries no information on dH)
            if h298[list(p)].sum() > h298[list(r)].sum():
                r, p = p, r
            sig = (tuple(sorted(r)), tuple(sorted(p)), kind)
            if sig not in seen and (sig[1], sig[0], kind) not in seen:
                seen.add(sig)
                break
        rs, ps = to_side(r), to_side(p)
        order = sum(n for _, n in rs)
        if kind == "el":
            a, b, e = arrhenius(order)
        elif kind == "tb":
            a, b, e = arrhenius(order + 1)
            sp, ef = efficiencies()
            tb_rxn.append(j)
            tb_sp.append(sp)
            tb_eff.append(ef)
        else:
            a, b, e = arrhenius(order + 1)           # LOW-pressure k0 (direct)
            ai, bi, ei = arrhenius(order)            # HIGH-pressure k_inf
            sp, ef = efficiencies()
            fo_rxn.append(j)
            fo_inf.append((ai, bi, ei))
            t2 = _loguniform(rng, 1e3, 1e4) if rng.random() < 0.5 else 0.0
            fo_troe.append((rng.uniform(0.1, 0.9), _loguniform(rng, 30.0, 1000.0),
                            _loguniform(rng, 500.0, 5000.0), t2))
            fo_sp.append(sp)
            fo_eff.append(ef)
        lnA.append(a)
        beta.append(b)
        EoR.append(e)
        flags.append(RXN_REVERSIBLE)
        reac.append(rs)
        prod.append(ps)

    def csr(lists):
        ptr = np.zeros(len(lists) + 1, dtype=np.int32)
        ptr[1:] = np.cumsum([len(x) for x in lists])
        return ptr

    reac_ptr, prod_ptr = csr(reac), csr(prod)
    nfo = len(fo_rxn)
    return Mechanism(
        species=names, elements=atoms, nasa=nasa,
        lnA=np.array(lnA), beta=np.array(beta), E_over_R=np.array(EoR), rflags=np.array(flags),
        reac_ptr=reac_ptr, reac_sp=np.array([s for x in reac for s, _ in x]),
        reac_nu=np.array([float(n) for x in reac for _, n in x]),
        prod_ptr=prod_ptr, prod_sp=np.array([s for x in prod for s, _ in x]),
        prod_nu=np.array([float(n) for x in prod for _, n in x]),
        tb_rxn=np.array(tb_rxn, dtype=np.int32), tb_ptr=csr(tb_sp),
        tb_sp=np.array([s for x in tb_sp for s in x], dtype=np.int32),
        tb_eff_m1=np.array([e for x in tb_eff for e in x]),
        fo_rxn=np.array(fo_rxn, dtype=np.int32),
        fo_type=np.full(nfo, FALLOFF_TROE, dtype=np.int32),
        fo_inf=np.array(fo_inf).reshape(nfo, 3), fo_troe=np.array(fo_troe).reshape(nfo, 4),
        fo_single=np.full(nfo, -1, dtype=np.int32), fo_ptr=csr(fo_sp),
        fo_sp=np.array([s for x in fo_sp for s in x], dtype=np.int32),
        fo_eff_m1=np.array([e for x in fo_eff for e in x]),
        name=name or f"synthetic-{ns}sp-{nr}r-seed{seed}",
    )


def custom_mechanism(species, nasa, reactions, elements=None, name="custom") -> Mechanism:
    """Small hand-built mechanisms for closed-form tests.

    ``reactions`` is a list of dicts with keys ``reac`` / ``prod`` (lists of
    (species index, coefficient)), ``lnA``, ``beta``, ``E_over_R``, optional
    ``reversible`` (default True), ``rev`` (explicit reverse (lnA, beta, EoR)),
    ``tb`` (list of (species, eff)), ``falloff`` (dict with ``inf`` = (lnA, beta,
    EoR), ``troe`` = (a, T3, T1, T2) or None, ``sri`` = (a, b, c, d, e) or None,
    ``cabr`` = True for a chemically activated reaction, ``single`` = species or
    -1, ``eff`` = list of (species, eff)), ``plog`` (list of (P [Pa], lnA,
    beta, EoR) levels) and ``cheb`` (dict with ``a`` = N x M coefficients, ``T`` =
    (Tmin, Tmax), ``P`` = (Pmin, Pmax) [Pa], ``conv`` = multiplier).
    """
    ns = len(species)
    nasa = np.asarray(nasa, dtype=np.float64).reshape(ns, 18)
    if elements is None:
        elements = np.zeros((ns, 4))
    lnA, beta, EoR, flags, reac, prod = [], [], [], [], [], []
    tb_rxn, tb_sp, tb_eff = [], [], []
    fo_rxn, fo_type, fo_inf, fo_troe, fo_single, fo_sp, fo_eff = [], [], [], [], [], [], []
    fo_sri, rev_arr, pl_rxn, pl_lev = [], np.zeros((len(reactions), 3)), [], []
    ch_rxn, ch_dims, ch_lim, ch_a = [], [], [], []
    for j, r in enumerate(reactions):
        lnA.append(r["lnA"])
        beta.append(r.get("beta", 0.0))
        EoR.append(r.get("E_over_R", 0.0))
        fl = RXN_REVERSIBLE if r.get("reversible", True) else 0
        if r.get("negative_A", False):
            fl |= RXN_NEGATIVE_A
        if r.get("rev") is not None:
            fl |= RXN_REVERSIBLE | RXN_EXPLICIT_REV
            rev_arr[j] = r["rev"]
        if r.get("plog") is not None:
            pl_rxn.append(j)
            pl_lev.append([tuple(v) for v in r["plog"]])
        if r.get("cheb") is not None:
            cb = r["cheb"]
            a = np.atleast_2d(np.asarray(cb["a"], dtype=np.float64))
            ch_rxn.append(j)
            ch_dims.append(a.shape)
            ch_lim.append((*cb["T"], *cb["P"], cb.get("conv", 1.0)))
            ch_a.append(a.ravel().tolist())
        flags.append(fl)
        reac.append(list(r["reac"]))
        prod.append(list(r["prod"]))
        if "tb" in r:
            tb_rxn.append(j)
            tb_sp.append([s for s, _ in r["tb"]])
            tb_eff.append([e - 1.0 for _, e in r["tb"]])
        if "falloff" in r:
            f = r["falloff"]
            fo_rxn.append(j)
            form = (FALLOFF_TROE if f.get("troe") is not None else
                    FALLOFF_SRI if f.get("sri") is not None else FALLOFF_LINDEMANN)
            fo_type.append(form | (FALLOFF_CHEMACT if f.get("cabr") else 0))
            fo_sri.append(f.get("sri") or (0.0, 0.0, 1.0, 1.0, 0.0))
            fo_inf.append(f["inf"])
            fo_troe.append(f.get("troe") or (0.0, 1.0, 1.0, 0.0))
            fo_single.append(f.get("single", -1))
            fo_sp.append([s for s, _ in f.get("eff", [])])
            fo_eff.append([e - 1.0 for _, e in f.get("eff", [])])

    def csr(lists):
        ptr = np.zeros(len(lists) + 1, dtype=np.int32)
        ptr[1:] = np.cumsum([len(x) for x in lists])
        return ptr

    nfo = len(fo_rxn)
    return Mechanism(
        species=list(species), elements=np.asarray(elements, dtype=np.float64), nasa=nasa,
        lnA=np.array(lnA, dtype=np.float64), beta=np.array(beta, dtype=np.float64),
        E_over_R=np.array(EoR, dtype=np.float64), rflags=np.array(flags),
        reac_ptr=csr(reac), reac_sp=np.array([s for x in reac for s, _ in x]),
        reac_nu=np.array([float(n) for x in reac for _, n in x]),
        prod_ptr=csr(prod), prod_sp=np.array([s for x in prod for s, _ in x]),
        prod_nu=np.array([float(n) for x in prod for _, n in x]),
        tb_rxn=np.array(tb_rxn, dtype=np.int32), tb_ptr=csr(tb_sp),
        tb_sp=np.array([s for x in tb_sp for s in x], dtype=np.int32),
        tb_eff_m1=np.array([e for x in tb_eff for e in x], dtype=np.float64),
        fo_rxn=np.array(fo_rxn, dtype=np.int32), fo_type=np.array(fo_type, dtype=np.int32),
        fo_inf=np.array(fo_inf, dtype=np.float64).reshape(nfo, 3),
        fo_troe=np.array(fo_troe, dtype=np.float64).reshape(nfo, 4),
        fo_single=np.array(fo_single, dtype=np.int32), fo_ptr=csr(fo_sp),
        fo_sp=np.array([s for x in fo_sp for s in x], dtype=np.int32),
        fo_eff_m1=np.array([e for x in fo_eff for e in x], dtype=np.float64),
        fo_sri=np.array(fo_sri, dtype=np.float64).reshape(nfo, 5),
        rev_arr=rev_arr, pl_rxn=np.array(pl_rxn, dtype=np.int32), pl_ptr=csr(pl_lev),
        pl_par=np.array([v for x in pl_lev for v in x], dtype=np.float64).reshape(-1, 4),
        ch_rxn=np.array(ch_rxn, dtype=np.int32), ch_dims=np.array(ch_dims, dtype=np.int32).reshape(-1, 2),
        ch_lim=np.array(ch_lim, dtype=np.float64).reshape(-1, 5), ch_ptr=csr(ch_a),
        ch_a=np.array([v for x in ch_a for v in x], dtype=np.float64),
        name=name,
    )


def constant_cp_nasa(cp_over_R, h298_over_R, s298_over_R, mw, tmid=1000.0):
    """NASA-7 row for a constant cp/R (identical low and high sets)."""
    T0 = 298.15
    a = [cp_over_R, 0.0, 0.0, 0.0, 0.0]
    a6 = h298_over_R - cp_over_R * T0
    a7 = s298_over_R - cp_over_R * np.log(T0)
    return a + [a6, a7] + a + [a6, a7] + [200.0, 6000.0, tmid, mw]


# The synthetic mechanisms of the five workloads (seeded, element balanced).
MECHANISMS = {
    14: dict(ns=14, nr=34, n_tb=4, n_troe=2, seed=1401),
    53: dict(ns=53, nr=325, n_tb=25, n_troe=12, seed=5301),
    160: dict(ns=160, nr=1000, n_tb=77, n_troe=37, seed=16001),
}

_CACHE = {}


def workload_mechanism(ns: int) -> Mechanism:
    """The seeded 14-, 53- or 160-species synthetic mechanism."""
    if ns not in _CACHE:
        _CACHE[ns] = synthetic_mechanism(**MECHANISMS[ns])
    return _CACHE[ns]
