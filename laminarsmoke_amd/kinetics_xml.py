"""OpenSMOKE++ ``kinetics.xml`` (the preprocessed mechanism a laminarSMOKE case loads)
<-> Mechanism tables.

Reader: the sections read by
  * ThermodynamicsMap::ImportElementsFromXMLFile (maps/ThermodynamicsMap.hpp:84-116):
    NumberOfElements, NamesOfElements, AtomicComposition (optional here);
  * ThermodynamicsMap_CHEMKIN::ImportSpeciesFromXMLFile / ImportCoefficientsFromXMLFile
    (maps/ThermodynamicsMap_CHEMKIN.hpp:359-410): NumberOfSpecies, NamesOfSpecies,
    Thermodynamics type="NASA" / NASA-coefficients, 18 numbers per species in the
    order of SetCoefficients (:263-343), which is the order of Mechanism.nasa;
  * KineticsMap_CHEMKIN::ImportCoefficientsFromXMLFile
    (maps/KineticsMap_CHEMKIN.hpp:311-1165): Kinetics type="OpenSMOKE"
    version="04-22-2013", NumberOfReactions, the 1-based index lists (Irreversible,
    Reversible, Reversible-Thermodynamics, Reversible-Explicit, ThreeBody, FallOff
    with Lindemann/Troe/SRI, CABR, ...), KineticParameters (Direct lnA /
    negative-lnA / Beta / E_over_R; ThreeBody per-reaction species and efficiency
    lists; FallOff and CABR k_inf lnA / Beta / E_over_R and per-reaction "lindemann" |
    "troe n a b c [d]" | "sri n a b c [d e]" followed by "species k" or "thirdbody"
    and the species / efficiency lists; Reverse lnA / Beta / E_over_R of the
    Reversible-Explicit reactions; PressureLog per-reaction "n (P A beta E)... convA
    convE" lists, P in atm (kernel/kinetics/PressureLogarithmicRateExpression.hpp:71-102,
    121-132); Chebyshev per-reaction "n N M a11 .. aNM conv", "Pmin Pmax" [atm],
    "Tmin Tmax" (kernel/kinetics/ChebyshevPolynomialRateExpression.hpp:71-93, 130-149));
  * StoichiometricMap::ReadFromASCIIFile (maps/StoichiometricMap.hpp:78-254): the
    species-major lists by stoichiometric class (1, 2, 3, 0.5, other) for reactants
    (numDir/jDir/valueDir5), products of all reactions (numRevTot/jRevTot/
    valueRevTot5) and of reversible ones (numRevEq/...), changeOfMoles and the
    explicit-reaction-order flag.
Every list is "n v1 ... vn" (OpenSMOKEVector::Load, math/OpenSMOKEVector.hpp:509-520).

Writer: the same layout as the reference preprocessor's
(preprocessing/PreProcessorKineticsPolicy_CHEMKIN.hpp:1877-2330,
preprocessing/PreProcessorSpecies.hpp:362-392,
kernel/kinetics/ReactionPolicy_CHEMKIN.hpp:2584-2620), so that tables read from
CHEMKIN input can be stored the way laminarSMOKE cases keep them.

Explicit forward reaction orders (the lambda lists after the flag) are read into
Mechanism.ord_*; explicit reverse orders that differ from the product coefficients,
and forms the kernels do not evaluate (extended PLOG and fall-off, FIT1, Janev-Langer,
Landau-Teller), raise KineticsXmlError.
"""
from __future__ import annotations

import xml.etree.ElementTree as ET

import numpy as np

from .chemkin import ATOMIC_WEIGHTS
from .mechanism import (FALLOFF_CHEMACT, FALLOFF_LINDEMANN, FALLOFF_SRI, FALLOFF_TROE, R_J_KMOL, RXN_EXPLICIT_REV,
                        RXN_NEGATIVE_A, RXN_REVERSIBLE, Mechanism)

VERSION = "04-22-2013"
_UNSUPPORTED_LISTS = ("ExtendedPressureLog", "ExtendedFallOff", "FIT1", "JanevLanger", "LandauTeller")


class KineticsXmlError(ValueError):
    pass


# ----------------------------------------------------------------------------------
# reading
# ----------------------------------------------------------------------------------
class _Tokens:
    def __init__(self, text):
        self.t = (text or "").split()
        self.i = 0

    def word(self):
        if self.i >= len(self.t):
            raise KineticsXmlError("unexpected end of a list")
        self.i += 1
        return self.t[self.i - 1]

    def vec(self, conv=float):
        n = int(self.word())
        return [conv(self.word()) for _ in range(n)]

    def more(self):
        return self.i < len(self.t)


def _node(parent, tag, required=True):
    n = parent.find(tag)
    if n is None and required:
        raise KineticsXmlError(f"<{tag}> not found")
    return n


def _ilist(parent, tag, required=True):
    n = _node(parent, tag, required)
    return [] if n is None else _Tokens(n.text).vec(int)


def read_kinetics_xml(text: str, name: str = "kinetics.xml") -> Mechanism:
    root = ET.fromstring(text)
    if root.tag != "opensmoke":
        root = _node(root, "opensmoke")
    ns = int(_node(root, "NumberOfSpecies").text.split()[0])
    species = _node(root, "NamesOfSpecies").text.split()[:ns]
    if len(species) != ns:
        raise KineticsXmlError("NamesOfSpecies is shorter than NumberOfSpecies")

    th = _node(root, "Thermodynamics")
    if th.get("type") != "NASA":
        raise KineticsXmlError("Thermodynamics type not supported")
    coef = np.array(_node(th, "NASA-coefficients").text.split(), dtype=np.float64)
    if coef.size < 18 * ns:
        raise KineticsXmlError("NASA-coefficients: fewer than 18 numbers per species")
    nasa = coef[:18 * ns].reshape(ns, 18)

    el_node = root.find("NumberOfElements")
    if el_node is not None:
        ne_ = int(el_node.text.split()[0])
        el_names = tuple(_node(root, "NamesOfElements").text.split()[:ne_])
        atoms = np.array(_node(root, "AtomicComposition").text.split(), dtype=np.float64)[:ns * ne_].reshape(ns, ne_)
        aw = np.array([ATOMIC_WEIGHTS.get(e.upper(), 0.0) for e in el_names])
    else:
        el_names, atoms, aw = ("C", "H", "O", "N"), np.zeros((ns, 4)), None

    kin = _node(root, "Kinetics")
    if kin.get("type") != "OpenSMOKE" or kin.get("version") != VERSION:
        raise KineticsXmlError("The current kinetic scheme is not supported (type/version)")
    nr = int(_node(kin, "NumberOfReactions").text.split()[0])
    reversible = set(_ilist(kin, "Reversible"))
    explicit = _ilist(kin, "Reversible-Explicit", required=False)
    if not set(explicit) <= reversible:
        raise KineticsXmlError("Reversible-Explicit reactions must be in the Reversible list")
    for tag in _UNSUPPORTED_LISTS:
        n = kin.find(tag)
        if n is not None and _Tokens(n.text).vec(int):
            raise KineticsXmlError(f"{tag} reactions are not supported")
    tb_idx = _ilist(kin, "ThreeBody")
    fo_node = _node(kin, "FallOff")
    fo_idx = _Tokens(fo_node.text).vec(int)
    cabr_node = kin.find("CABR")
    cabr_idx = _Tokens(cabr_node.text).vec(int) if cabr_node is not None else []
    pl_idx = _ilist(kin, "PressureLog", required=False)
    ch_idx = _ilist(kin, "Chebyshev", required=False)

    kp = _node(kin, "KineticParameters")
    direct = _node(kp, "Direct")
    lnA = np.array(_Tokens(_node(direct, "lnA").text).vec())
    neg = set(_ilist(direct, "negative-lnA", required=False))
    beta = np.array(_Tokens(_node(direct, "Beta").text).vec())
    EoR = np.array(_Tokens(_node(direct, "E_over_R").text).vec())
    if not (len(lnA) == len(beta) == len(EoR) == nr):
        raise KineticsXmlError("Direct parameters do not have NumberOfReactions entries")
    flags = np.zeros(nr, dtype=np.int32)
    for j in range(1, nr + 1):
        if j in reversible:
            flags[j - 1] |= RXN_REVERSIBLE
        if j in neg:
            flags[j - 1] |= RXN_NEGATIVE_A
    rev_arr = np.zeros((nr, 3))
    if explicit:                                   # KineticsMap_CHEMKIN.hpp:611-639
        rv = _node(kp, "Reverse")
        ra = [_Tokens(_node(rv, tag).text).vec() for tag in ("lnA", "Beta", "E_over_R")]
        if not all(len(v) == len(explicit) for v in ra):
            raise KineticsXmlError("Reverse parameters do not match the Reversible-Explicit list")
        for q, j in enumerate(explicit):
            flags[j - 1] |= RXN_EXPLICIT_REV
            rev_arr[j - 1] = (ra[0][q], ra[1][q], ra[2][q])

    tb_sp, tb_eff = [], []
    if tb_idx:
        nodes = list(_node(kp, "ThreeBody").findall("reaction"))
        if len(nodes) != len(tb_idx):
            raise KineticsXmlError("ThreeBody parameters do not match the ThreeBody list")
        for nd in nodes:
            t = _Tokens(nd.text)
            sp = [k - 1 for k in t.vec(int)]
            ef = [e - 1.0 for e in t.vec()]
            tb_sp.append(sp)
            tb_eff.append(ef)

    fo_type, fo_inf, fo_troe, fo_sri, fo_single, fo_sp, fo_eff = [], [], [], [], [], [], []
    # fall-off entries, then the chemically activated ones (KineticsMap_CHEMKIN.hpp:786-912)
    for tag, idx_list, chem in (("FallOff", fo_idx, 0), ("CABR", cabr_idx, FALLOFF_CHEMACT)):
        if not idx_list:
            continue
        fk = _node(kp, tag)
        a = _Tokens(_node(fk, "lnA").text).vec()
        b = _Tokens(_node(fk, "Beta").text).vec()
        e = _Tokens(_node(fk, "E_over_R").text).vec()
        nodes = list(_node(fk, "Parameters").findall("reaction"))
        if not (len(a) == len(b) == len(e) == len(nodes) == len(idx_list)):
            raise KineticsXmlError(f"{tag} parameters do not match the {tag} list")
        for k, nd in enumerate(nodes):
            fo_inf.append((a[k], b[k], e[k]))
            t = _Tokens(nd.text)
            kind = t.word()
            troe, sri = (0.0, 1.0, 1.0, 0.0), (0.0, 0.0, 1.0, 1.0, 0.0)
            if kind == "lindemann":
                fo_type.append(FALLOFF_LINDEMANN | chem)
            elif kind == "troe":
                c = t.vec()
                if len(c) not in (3, 4):
                    raise KineticsXmlError("troe needs 3 or 4 parameters")
                troe = tuple(c) + ((0.0,) if len(c) == 3 else ())          # KineticsMap_CHEMKIN.hpp:736-746
                fo_type.append(FALLOFF_TROE | chem)
            elif kind == "sri":
                c = t.vec()
                if len(c) not in (3, 5):
                    raise KineticsXmlError("sri needs 3 or 5 parameters")
                sri = tuple(c) + ((1.0, 0.0) if len(c) == 3 else ())       # KineticsMap_CHEMKIN.hpp:749-767
                fo_type.append(FALLOFF_SRI | chem)
            else:
                raise KineticsXmlError(f"unknown fall-off form {kind!r}")
            fo_troe.append(troe)
            fo_sri.append(sri)
            how = t.word()
            if how == "species":
                fo_single.append(int(t.word()) - 1)
                fo_sp.append([])
                fo_eff.append([])
            else:
                fo_single.append(-1)
                fo_sp.append([k - 1 for k in t.vec(int)])
                fo_eff.append([v - 1.0 for v in t.vec()])

    # one fall-off table in reaction order (the FallOff and CABR lists are each sorted)
    fo_all = list(fo_idx) + list(cabr_idx)
    perm = sorted(range(len(fo_all)), key=lambda i: fo_all[i])
    fo_all = [fo_all[i] for i in perm]
    fo_type, fo_inf, fo_troe, fo_sri, fo_single, fo_sp, fo_eff = (
        [v[i] for i in perm] for v in (fo_type, fo_inf, fo_troe, fo_sri, fo_single, fo_sp, fo_eff))

    pl_lev = []
    if pl_idx:                                     # KineticsMap_CHEMKIN.hpp:927-937
        t = _Tokens(_node(kp, "PressureLog").text)
        for _ in pl_idx:
            c = t.vec()
            nlev = (len(c) - 2) // 4
            if nlev < 1 or len(c) != 4 * nlev + 2:
                raise KineticsXmlError("PressureLog entry is not n (P A beta E)... convA convE")
            ca, ce = c[-2], c[-1]
            lev = []
            for i in range(nlev):
                P, A, bp, E = c[4 * i:4 * i + 4]
                if A < 0.0:
                    raise KineticsXmlError("PLOG needs non-negative frequency factors")
                lev.append((P * 101325.0, np.log(max(A * ca, 1e-32)), bp, E * ce / R_J_KMOL))
            if any(y[0] <= x[0] for x, y in zip(lev[:-1], lev[1:])):
                raise KineticsXmlError("PLOG pressures must increase")
            pl_lev.append(lev)

    ch_dims, ch_lim, ch_a = [], [], []
    if ch_idx:                                     # KineticsMap_CHEMKIN.hpp:915-925
        t = _Tokens(_node(kp, "Chebyshev").text)
        for _ in ch_idx:
            c = t.vec()
            if len(c) < 3 or int(c[0]) < 1 or int(c[1]) < 1 or int(c[0]) * int(c[1]) != len(c) - 3:
                raise KineticsXmlError("Chebyshev entry is not n N M a11 .. aNM conv")
            pmin, pmax, tmin, tmax = (float(t.word()) for _ in range(4))
            ch_dims.append((int(c[0]), int(c[1])))
            ch_lim.append((tmin, tmax, pmin * 101325.0, pmax * 101325.0, c[-1]))
            ch_a.append(c[2:-1])

    st = _node(kin, "Stoichiometry")
    if st.get("type") != "OpenSMOKE" or st.get("version") != VERSION:
        raise KineticsXmlError("The current stoichiometric data are not supported")
    reac, prod, orders = _read_stoichiometry(_Tokens(st.text), ns, nr, reversible)

    def csr(lists):
        ptr = np.zeros(len(lists) + 1, dtype=np.int32)
        ptr[1:] = np.cumsum([len(x) for x in lists])
        return ptr

    nfo = len(fo_idx) + len(cabr_idx)
    return Mechanism(
        species=list(species), elements=atoms, nasa=nasa,
        lnA=lnA, beta=beta, E_over_R=EoR, rflags=flags,
        reac_ptr=csr(reac), reac_sp=np.array([s for x in reac for s, _ in x], dtype=np.int32),
        reac_nu=np.array([v for x in reac for _, v in x], dtype=np.float64),
        prod_ptr=csr(prod), prod_sp=np.array([s for x in prod for s, _ in x], dtype=np.int32),
        prod_nu=np.array([v for x in prod for _, v in x], dtype=np.float64),
        tb_rxn=np.array([j - 1 for j in tb_idx], dtype=np.int32), tb_ptr=csr(tb_sp),
        tb_sp=np.array([s for x in tb_sp for s in x], dtype=np.int32),
        tb_eff_m1=np.array([v for x in tb_eff for v in x], dtype=np.float64),
        fo_rxn=np.array([j - 1 for j in fo_all], dtype=np.int32),
        fo_type=np.array(fo_type, dtype=np.int32),
        fo_inf=np.array(fo_inf, dtype=np.float64).reshape(nfo, 3),
        fo_troe=np.array(fo_troe, dtype=np.float64).reshape(nfo, 4),
        fo_single=np.array(fo_single, dtype=np.int32), fo_ptr=csr(fo_sp),
        fo_sp=np.array([s for x in fo_sp for s in x], dtype=np.int32),
        fo_eff_m1=np.array([v for x in fo_eff for v in x], dtype=np.float64),
        fo_sri=np.array(fo_sri, dtype=np.float64).reshape(nfo, 5),
        ord_ptr=csr(orders), ord_sp=np.array([s for x in orders for s, _ in x], dtype=np.int32),
        ord_val=np.array([v for x in orders for _, v in x], dtype=np.float64),
        rev_arr=rev_arr, pl_rxn=np.array([j - 1 for j in pl_idx], dtype=np.int32),
        pl_ptr=np.concatenate([[0], np.cumsum([len(x) for x in pl_lev])]).astype(np.int32),
        pl_par=np.array([v for x in pl_lev for v in x], dtype=np.float64).reshape(-1, 4),
        ch_rxn=np.array([j - 1 for j in ch_idx], dtype=np.int32),
        ch_dims=np.array(ch_dims, dtype=np.int32).reshape(-1, 2),
        ch_lim=np.array(ch_lim, dtype=np.float64).reshape(-1, 5), ch_ptr=csr(ch_a),
        ch_a=np.array([v for x in ch_a for v in x], dtype=np.float64),
        name=name, element_names=tuple(e.upper() for e in el_names), atomic_weights=aw,
    )


_CLASS_NU = (1.0, 2.0, 3.0, 0.5)


def _read_stoichiometry(t: _Tokens, ns: int, nr: int, reversible):
    """Reaction-major (species, nu) lists from StoichiometricMap's species-major ASCII form,
    and the explicit forward orders of the reactions whose orders differ from their
    reactant coefficients."""
    num = {side: [t.vec(int) for _ in range(5)] for side in ("dir", "revtot", "reveq")}
    lists = {}
    for side in ("dir", "revtot", "reveq"):
        js = [t.vec(int) for _ in range(5)]
        v5 = t.vec()
        lists[side] = (js, v5)
    t.vec()                                   # changeOfMoles (recomputed from the stoichiometry)
    explicit = t.word() if t.more() else "0"
    if explicit not in ("0", "false", "1", "true"):
        raise KineticsXmlError("bad explicit-reaction-order flag")
    if explicit in ("1", "true"):             # lambda lists (StoichiometricMap.hpp:194-247)
        num["ldir"] = [t.vec(int) for _ in range(5)]
        num["lrev"] = [t.vec(int) for _ in range(5)]
        lists["ldir"] = ([t.vec(int) for _ in range(5)], t.vec())
        lists["lrev"] = ([t.vec(int) for _ in range(5)], t.vec())
    out = {}
    for side in [s for s in ("dir", "revtot", "reveq", "ldir", "lrev") if s in lists]:
        per = [dict() for _ in range(nr)]
        js, v5 = lists[side]
        for c in range(5):
            counts = num[side][c]
            if len(counts) != ns:
                raise KineticsXmlError("stoichiometry counts do not have NumberOfSpecies entries")
            q = 0
            for k in range(ns):
                for _ in range(counts[k]):
                    j = js[c][q] - 1
                    nu = _CLASS_NU[c] if c < 4 else v5[q]
                    if not (0 <= j < nr):
                        raise KineticsXmlError("stoichiometry reaction index out of range")
                    if k in per[j]:
                        raise KineticsXmlError(f"species {k + 1} listed twice on one side of reaction {j + 1}")
                    per[j][k] = nu
                    q += 1
            if q != len(js[c]):
                raise KineticsXmlError("stoichiometry index list length does not match its counts")
        out[side] = [sorted(d.items()) for d in per]
    orders = [[] for _ in range(nr)]
    if "ldir" in out:
        for j in range(nr):
            if out["ldir"][j] != out["dir"][j]:
                orders[j] = out["ldir"][j]
            rev = out["reveq"][j] if (j + 1) in reversible else []
            if (j + 1) in reversible and out["lrev"][j] != rev:
                raise KineticsXmlError(f"explicit reverse orders (RORD) of reaction {j + 1} are not supported")
    return out["dir"], out["revtot"], orders


def load_kinetics_xml(path: str) -> Mechanism:
    with open(path) as f:
        return read_kinetics_xml(f.read(), name=path)


# ----------------------------------------------------------------------------------
# writing
# ----------------------------------------------------------------------------------
def _vec(values, fmt="{:.16e}", per_line=1):
    vals = list(values)
    out = [str(len(vals))]
    for i in range(0, len(vals), per_line):
        out.append(" ".join(fmt.format(v) for v in vals[i:i + per_line]))
    return "\n".join(out)


def _ivec(values):
    vals = [int(v) for v in values]
    lines = [str(len(vals))]
    for i in range(0, len(vals), 30):
        lines.append(" ".join(str(v) for v in vals[i:i + 30]))
    return "\n".join(lines)


def _tag(tag, body, attrs=""):
    return f"<{tag}{attrs}>\n{body}\n</{tag}>"


def _stoichiometry(m: Mechanism) -> str:
    ns, nr = m.ns, m.nr
    rev = (m.rflags & RXN_REVERSIBLE) != 0

    def classify(ptr, sp, nu, only=None):
        entries = [[[] for _ in range(ns)] for _ in range(5)]       # class -> species -> [(j, nu)]
        for j in range(nr):
            if only is not None and not only[j]:
                continue
            for p in range(ptr[j], ptr[j + 1]):
                v = float(nu[p])
                c = _CLASS_NU.index(v) if v in _CLASS_NU else 4
                entries[c][int(sp[p])].append((j + 1, v))
        nums = [[len(entries[c][k]) for k in range(ns)] for c in range(5)]
        js = [[j for k in range(ns) for j, _ in entries[c][k]] for c in range(5)]
        v5 = [v for k in range(ns) for _, v in entries[4][k]]
        return nums, js, v5

    d = classify(m.reac_ptr, m.reac_sp, m.reac_nu)
    t = classify(m.prod_ptr, m.prod_sp, m.prod_nu)
    e = classify(m.prod_ptr, m.prod_sp, m.prod_nu, only=rev)
    parts = [_ivec(x) for x in d[0]] + [_ivec(x) for x in t[0]] + [_ivec(x) for x in e[0]]
    for blk in (d, t, e):
        parts += [_ivec(x) for x in blk[1]] + [_vec(blk[2])]
    dn = [float(m.prod_nu[m.prod_ptr[j]:m.prod_ptr[j + 1]].sum() - m.reac_nu[m.reac_ptr[j]:m.reac_ptr[j + 1]].sum())
          for j in range(nr)]
    parts += [_vec(dn)]
    if m.ord_ptr[-1] == 0:
        parts.append("0")
    else:
        # lambda lists: forward orders (explicit where given, else the coefficients), reverse
        # orders = the product coefficients of reversible reactions
        optr, osp, oval = [0], [], []
        for j in range(nr):
            a, b = (m.ord_ptr[j], m.ord_ptr[j + 1]) if m.ord_ptr[j + 1] > m.ord_ptr[j] else (None, None)
            if a is None:
                osp += m.reac_sp[m.reac_ptr[j]:m.reac_ptr[j + 1]].tolist()
                oval += m.reac_nu[m.reac_ptr[j]:m.reac_ptr[j + 1]].tolist()
            else:
                osp += m.ord_sp[a:b].tolist()
                oval += m.ord_val[a:b].tolist()
            optr.append(len(osp))
        ld = classify(np.array(optr), np.array(osp), np.array(oval))
        parts.append("1")
        parts += [_ivec(x) for x in ld[0]] + [_ivec(x) for x in e[0]]
        parts += [_ivec(x) for x in ld[1]] + [_vec(ld[2])] + [_ivec(x) for x in e[1]] + [_vec(e[2])]
    return "\n".join(parts)


def write_kinetics_xml(m: Mechanism) -> str:
    ns, nr = m.ns, m.nr
    idx = lambda mask: [j + 1 for j in range(nr) if mask[j]]
    rev = (m.rflags & RXN_REVERSIBLE) != 0
    fo_of = {int(j): k for k, j in enumerate(m.fo_rxn)}
    lines = ['<?xml version="1.0" encoding="utf-8"?>', '<opensmoke version="0.1a">']
    lines.append(_tag("NumberOfElements", str(len(m.element_names))))
    lines.append(_tag("NamesOfElements", " ".join(m.element_names)))
    lines.append(_tag("AtomicComposition", "\n".join(" ".join(f"{v:.10g}" for v in row) for row in m.elements)))
    lines.append(_tag("NumberOfSpecies", str(ns)))
    lines.append(_tag("NamesOfSpecies", " ".join(m.species)))
    lines.append(_tag("Thermodynamics", _tag("NASA-coefficients",
                                             "\n".join(" ".join(f"{v:.16e}" for v in row) for row in m.nasa)),
                      ' type="NASA"'))
    k = [f'<Kinetics type="OpenSMOKE" version="{VERSION}">', _tag("NumberOfReactions", str(nr))]
    k.append(_tag("Irreversible", _ivec(idx(~rev))))
    xrev = (m.rflags & RXN_EXPLICIT_REV) != 0
    k.append(_tag("Reversible", _ivec(idx(rev))))
    k.append(_tag("Reversible-Thermodynamics", _ivec(idx(rev & ~xrev))))
    k.append(_tag("Reversible-Explicit", _ivec(idx(xrev))))
    k.append(_tag("ThreeBody", _ivec(m.tb_rxn + 1)))
    fo_types = m.fo_type.tolist()
    groups = {}
    for chem, tag in ((0, "FallOff"), (FALLOFF_CHEMACT, "CABR")):
        sel = [i for i, t in enumerate(fo_types) if (t & FALLOFF_CHEMACT) == chem]
        groups[tag] = sel
        by = lambda form: _ivec([int(m.fo_rxn[i]) + 1 for i in sel if fo_types[i] & 3 == form])
        k.append(f"<{tag}>\n" + _ivec([int(m.fo_rxn[i]) + 1 for i in sel]) + "\n" +
                 _tag("Lindemann", by(FALLOFF_LINDEMANN)) + "\n" + _tag("Troe", by(FALLOFF_TROE)) + "\n" +
                 _tag("SRI", by(FALLOFF_SRI)) + f"\n</{tag}>")
    if len(m.ch_rxn):
        k.append(_tag("Chebyshev", _ivec(m.ch_rxn + 1)))
    if len(m.pl_rxn):
        k.append(_tag("PressureLog", _ivec(m.pl_rxn + 1)))
    neg = [j + 1 for j in range(nr) if m.rflags[j] & RXN_NEGATIVE_A]
    direct = "\n".join([_tag("lnA", _vec(m.lnA)), _tag("negative-lnA", _ivec(neg)),
                        _tag("Beta", _vec(m.beta)), _tag("E_over_R", _vec(m.E_over_R))])
    kp = [_tag("Direct", direct)]
    if xrev.any():
        rj = np.nonzero(xrev)[0]
        kp.append(_tag("Reverse", "\n".join(_tag(t, _vec(m.rev_arr[rj, i]))
                                            for i, t in enumerate(("lnA", "Beta", "E_over_R")))))
    if len(m.tb_rxn):
        rs = []
        for s in range(len(m.tb_rxn)):
            a, b = m.tb_ptr[s], m.tb_ptr[s + 1]
            body = _ivec(m.tb_sp[a:b] + 1) + "\n" + _vec(m.tb_eff_m1[a:b] + 1.0, per_line=30)
            rs.append(_tag("reaction", body, f' i="{int(m.tb_rxn[s]) + 1}"'))
        kp.append(_tag("ThreeBody", "\n".join(rs)))
    for tag in ("FallOff", "CABR"):
        sel = groups[tag]
        if not sel:
            continue
        ps = []
        for s in sel:
            t = fo_types[s] & 3
            if t == FALLOFF_LINDEMANN:
                head = "lindemann"
            elif t == FALLOFF_TROE:
                tr = list(m.fo_troe[s])
                head = "troe " + " ".join([str(4 if tr[3] != 0.0 else 3)] + [f"{v:.16e}" for v in tr[:4 if tr[3] != 0.0 else 3]])
            else:
                head = "sri 5 " + " ".join(f"{v:.16e}" for v in m.fo_sri[s])
            if m.fo_single[s] >= 0:
                tail = f"species {int(m.fo_single[s]) + 1}"
            else:
                a, b = m.fo_ptr[s], m.fo_ptr[s + 1]
                tail = "thirdbody\n" + _ivec(m.fo_sp[a:b] + 1) + "\n" + _vec(m.fo_eff_m1[a:b] + 1.0, per_line=30)
            ps.append(_tag("reaction", head + "\n" + tail, f' i="{int(m.fo_rxn[s]) + 1}"'))
        inf = m.fo_inf[sel]
        kp.append(_tag(tag, "\n".join([_tag("lnA", _vec(inf[:, 0])), _tag("Beta", _vec(inf[:, 1])),
                                       _tag("E_over_R", _vec(inf[:, 2])), _tag("Parameters", "\n".join(ps))])))
    if len(m.ch_rxn):
        es = []
        for s in range(len(m.ch_rxn)):
            N, M = (int(v) for v in m.ch_dims[s])
            vals = [N, M] + m.ch_a[m.ch_ptr[s]:m.ch_ptr[s + 1]].tolist() + [m.ch_lim[s, 4]]
            Tl, Tu, Pl, Pu = m.ch_lim[s, :4]
            es.append(f"{len(vals)}\n" + " ".join(f"{v:.17g}" for v in vals) +
                      f"\n{Pl / 101325.0:.17g} {Pu / 101325.0:.17g}\n{Tl:.17g} {Tu:.17g}")
        kp.append(_tag("Chebyshev", "\n".join(es)))
    if len(m.pl_rxn):
        # raw values with unit factors 1 (A) and R (E): P [atm], A, beta, E/R
        es = []
        for s in range(len(m.pl_rxn)):
            lev = m.pl_par[m.pl_ptr[s]:m.pl_ptr[s + 1]]
            vals = [v for P, a, b, e in lev for v in (P / 101325.0, np.exp(a), b, e)] + [1.0, R_J_KMOL]
            es.append(f"{len(vals)}\n" + " ".join(f"{v:.17g}" for v in vals))
        kp.append(_tag("PressureLog", "\n".join(es)))
    k.append(_tag("KineticParameters", "\n".join(kp)))
    k.append(_tag("Stoichiometry", _stoichiometry(m), f' type="OpenSMOKE" version="{VERSION}"'))
    k.append("</Kinetics>")
    lines += k
    lines.append("</opensmoke>")
    return "\n".join(lines) + "\n"


def main(argv=None):
    """python -m laminarsmoke_amd.kinetics_xml KINETICS.CKI THERMO.CKT OUT.xml"""
    import sys
    from .chemkin import load_chemkin
    args = sys.argv[1:] if argv is None else argv
    if len(args) != 3:
        print(main.__doc__)
        return 2
    m = load_chemkin(args[0], args[1])
    with open(args[2], "w") as f:
        f.write(write_kinetics_xml(m))
    print(f"{args[2]}: {m.ns} species, {m.nr} reactions")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
