"""CHEMKIN-II gas-phase input (kinetics file + thermodynamics file) -> Mechanism tables.

This is the work the OpenSMOKE++ preprocessor does before a laminarSMOKE run
(kernel/thermo/ThermoPolicy_CHEMKIN.hpp, kernel/kinetics/ReactionPolicy_CHEMKIN.hpp),
folded into the flat SI tables of :class:`laminarsmoke_amd.mechanism.Mechanism`
that the maps hold after loading (maps/KineticsMap_CHEMKIN.hpp:311-1165).

Supported:

* ``ELEMENTS`` / ``SPECIES`` / ``REACTIONS [units]`` blocks, ``!`` comments;
* reversible (``=``, ``<=>``) and irreversible (``=>``) reactions, stoichiometric
  coefficients, repeated species on one side merged (``OH+OH`` -> 2 OH);
* third-body reactions (``+M``) with efficiencies;
* fall-off reactions (``(+M)`` or ``(+species)``) with ``LOW`` and optional
  3- or 4-parameter ``TROE`` or 3- or 5-parameter ``SRI``, with efficiencies;
* chemically activated reactions (``(+M)`` with ``HIGH`` instead of ``LOW``,
  with optional ``TROE`` / ``SRI``): the reaction line is k0 (order), ``HIGH`` is
  k_inf (order - 1) (ReactionPolicy_CHEMKIN.hpp:546-582, 1396-1407);
* explicit reverse Arrhenius parameters (``REV``) on reversible elementary and
  third-body reactions, converted with the product order (+1 for ``+M``)
  (ReactionPolicy_CHEMKIN.hpp:882-902, 1359-1376, 1474-1478);
* pressure-logarithmic reactions (``PLOG /P A beta E/`` lines, P in atm,
  increasing) on reactions without ``M``: A converted like the reaction line and
  floored at 1e-32 in the log (PressureLogarithmicRateExpression.hpp:71-102);
* Chebyshev reactions (``(+M)`` with ``CHEB/ N M a11 ... aNM /``, optional
  ``TCHEB/ Tmin Tmax /`` and ``PCHEB/ Pmin Pmax /`` in atm, defaults 300-2500 K and
  0.001-100 atm; k in the reaction's A units, converted with 1e3^(1 - order))
  (ReactionPolicy_CHEMKIN.hpp:456-474, 751-770, 1072-1092, 1322-1344, 1409-1412);
* ``DUPLICATE``; negative pre-exponential factors on elementary reactions;
* explicit forward reaction orders (``FORD``): the orders are the reactant
  coefficients with the named species' orders replaced or added
  (ReactionPolicy_CHEMKIN.hpp:511-545), and the A conversion uses their sum;
* energy units CAL/MOLE (default), KCAL/MOLE, JOULES/MOLE, KJOULES/MOLE, KELVINS;
  quantity unit MOLES.

Anything else the kernels do not evaluate (``RORD``, ``LT``, ``MOLECULES``, ...) raises
:class:`ChemkinError` naming the reaction, rather than being dropped.

Unit conversions follow ReactionPolicy_CHEMKIN::ConvertUnits /
ConversionFactors (ReactionPolicy_CHEMKIN.hpp:1378-1501): E [J/kmol] = E x
e_conversion (1e3 J_from_cal for CAL/MOLE, J_from_cal = 4.18443); A [m, kmol, s] =
A x 1e3^(1 - order), with order = sum of reactant coefficients (+1 for ``+M``);
for fall-off reactions the LOW set has order + 1 and becomes the direct
parameters, the Arrhenius line becomes k_inf (ReactionPolicy_CHEMKIN.hpp:618-631).
E_over_R = E / R with R = 8314.4621 J/kmol/K (math/PhysicalConstants.h:20).
Molecular weights use the atomic weights of kernel/thermo/AtomicElementMap.hpp:41-60.

Thermodynamics: the fixed-column NASA-7 format read by
ThermoPolicy_CHEMKIN::ReadThermodynamicsFromStrings (ThermoPolicy_CHEMKIN.hpp:85-330):
line 1 holds the name, up to 4 (+1 in columns 74-78) element/count pairs and
T_low, T_high, T_mid (defaults from the ``THERMO ALL`` line); lines 2-4 hold the
14 coefficients, which are already in the order of the 18-entry table
(high a1..a7, low a1..a7).
"""
from __future__ import annotations

import os
import re

import numpy as np

from .mechanism import (FALLOFF_CHEMACT, FALLOFF_LINDEMANN, FALLOFF_SRI, FALLOFF_TROE, RXN_EXPLICIT_REV,
                        RXN_NEGATIVE_A, RXN_REVERSIBLE, J_FROM_CAL, R_J_KMOL, Mechanism)

# kernel/thermo/AtomicElementMap.hpp:41-60 (the values are stored as floats there)
ATOMIC_WEIGHTS = {
    "H": 1.008000016212463, "HE": 4.002999782562256, "C": 12.010999679565430,
    "N": 14.0069999694824, "O": 15.998999595642090, "F": 18.99840, "NE": 20.18300,
    "S": 32.06, "CL": 35.45300, "AR": 39.948001861572270,
}

E_CONVERSION = {               # J/kmol per unit (ReactionPolicy_CHEMKIN.hpp:1484-1489)
    "CAL/MOLE": 1e3 * J_FROM_CAL, "KCAL/MOLE": 1e6 * J_FROM_CAL,
    "JOULES/MOLE": 1e3, "KJOULES/MOLE": 1e6, "KELVINS": R_J_KMOL,
}
_UNSUPPORTED = {"RORD", "LT", "RLT",
                "JAN", "FIT1", "TDEP", "EXCI", "MOME", "XSMI", "USRPROG", "LOWMX", "LOWSP", "TROEMX",
                "TROESP", "SRIMX", "SRISP", "PLOGMX", "PLOGSP", "COLLEFF", "UNITS"}


class ChemkinError(ValueError):
    pass


# ----------------------------------------------------------------------------------
# thermodynamics
# ----------------------------------------------------------------------------------
def _coeffs(line: str, n: int) -> list:
    return [float(line[15 * i:15 * (i + 1)].replace("D", "E").replace("d", "e")) for i in range(n)]


def read_thermo(text: str) -> dict:
    """NASA-7 entries of a CHEMKIN thermodynamics file.

    Returns {NAME: (coef[14], (T_low, T_high, T_mid), {ELEMENT: count})}.
    """
    lines = [ln.rstrip("\n\r") for ln in text.splitlines()]
    out = {}
    defaults = (300.0, 5000.0, 1000.0)
    i = 0
    while i < len(lines) and not lines[i].strip().upper().startswith("THERMO"):
        i += 1
    if i == len(lines):
        raise ChemkinError("no THERMO block")
    i += 1
    # global temperature line (THERMO ALL): T_low T_mid T_high
    while i < len(lines) and (not lines[i].strip() or lines[i].lstrip().startswith("!")):
        i += 1
    tok = lines[i].split("!")[0].split()
    try:
        t = [float(x) for x in tok[:3]]
        defaults = (t[0], t[2], t[1])
        i += 1
    except (ValueError, IndexError):
        pass
    while i < len(lines):
        ln = lines[i]
        s = ln.strip()
        if not s or s.startswith("!"):
            i += 1
            continue
        if s.upper().startswith("END"):
            break
        if len(ln) < 80 or ln[79] != "1":
            i += 1
            continue
        name = ln[:18].split()[0].upper()
        elements = {}
        fields = [ln[24 + 5 * k: 29 + 5 * k] for k in range(4)] + [ln[73:78]]
        for f in fields:
            sym, cnt = f[:2].strip().upper(), f[2:].strip()
            if sym and not sym[0].isdigit() and cnt:
                c = float(cnt)
                if c != 0.0:
                    elements[sym] = elements.get(sym, 0.0) + c
        tl, th, tm = defaults
        for lo, hi, slot in ((45, 55, 0), (55, 65, 1), (65, 73, 2)):
            v = ln[lo:hi].strip()
            if v:
                if slot == 0:
                    tl = float(v)
                elif slot == 1:
                    th = float(v)
                else:
                    tm = float(v)
        if not (tl <= tm <= th):
            raise ChemkinError(f"{name}: temperature limits out of order ({tl}, {tm}, {th})")
        try:
            c = _coeffs(lines[i + 1], 5) + _coeffs(lines[i + 2], 5) + _coeffs(lines[i + 3], 4)
        except (ValueError, IndexError) as e:
            raise ChemkinError(f"{name}: malformed NASA-7 coefficient lines") from e
        if name not in out:                    # first entry wins
            out[name] = (c, (tl, th, tm), elements)
        i += 4
    return out


# ----------------------------------------------------------------------------------
# kinetics
# ----------------------------------------------------------------------------------
_AUX = re.compile(r"([^\s/]+)\s*/([^/]*)/")


def _strip(line: str) -> str:
    return line.split("!")[0].rstrip()


def _blocks(text: str):
    """Yield (keyword, list of lines) for the ELEMENTS, SPECIES and REACTIONS blocks."""
    lines = [_strip(ln) for ln in text.splitlines()]
    i = 0
    while i < len(lines):
        s = lines[i].strip()
        if not s:
            i += 1
            continue
        head = s.split()[0].upper()
        key = {"ELEMENTS": "ELEMENTS", "ELEM": "ELEMENTS", "SPECIES": "SPECIES", "SPEC": "SPECIES",
               "REACTIONS": "REACTIONS", "REAC": "REACTIONS", "THERMO": "THERMO"}.get(head)
        if key is None:
            raise ChemkinError(f"unexpected line outside a block: {s!r}")
        header = s.split()[1:]
        body = []
        i += 1
        if key != "REACTIONS":
            # ELEMENTS/SPECIES may list names on the header line and close on the same line
            rest = [w for w in header]
            if rest and rest[-1].upper() == "END":
                yield key, rest[:-1], []
                continue
            body.extend(rest)
            while i < len(lines) and lines[i].strip().upper() != "END" and \
                    not lines[i].strip().upper().endswith(" END"):
                body.extend(lines[i].split())
                i += 1
            if i < len(lines) and lines[i].strip().upper().endswith(" END"):
                body.extend(lines[i].split()[:-1])
            i += 1
            yield key, body, []
        else:
            while i < len(lines) and lines[i].strip().upper() != "END":
                if lines[i].strip():
                    body.append(lines[i].strip())
                i += 1
            i += 1
            yield key, header, body


def _split_side(side: str, species_set: set, what: str):
    """'H+O2(+M)' -> ([(name, nu)], third_body, falloff_collider)."""
    third, collider = False, None
    m = re.search(r"\(\+\s*([^)]+)\)", side)
    if m:
        collider = m.group(1).strip().upper()
        side = side[:m.start()] + side[m.end():]
    terms = [t for t in side.split("+") if t]
    out = []
    for t in terms:
        t = t.strip().upper()
        if t == "M":
            third = True
            continue
        if t in species_set:
            out.append((t, 1.0))
            continue
        mm = re.match(r"^(\d+\.?\d*|\.\d+)(.+)$", t)
        if mm and mm.group(2) in species_set:
            out.append((mm.group(2), float(mm.group(1))))
            continue
        raise ChemkinError(f"{what}: unknown species {t!r}")
    merged = {}
    for n, nu in out:
        merged[n] = merged.get(n, 0.0) + nu
    return list(merged.items()), third, collider


def _parse_reaction_line(line: str, species_set: set, idx: int):
    tok = line.split()
    if len(tok) < 4:
        raise ChemkinError(f"reaction {idx + 1}: expected an equation and A, beta, E: {line!r}")
    try:
        A, beta, E = (float(x.replace("D", "E").replace("d", "e")) for x in tok[-3:])
    except ValueError as e:
        raise ChemkinError(f"reaction {idx + 1}: cannot read A, beta, E from {line!r}") from e
    eq = "".join(tok[:-3]).upper()
    what = f"reaction {idx + 1} ({eq})"
    if "<=>" in eq:
        lhs, rhs = eq.split("<=>")
        rev = True
    elif "=>" in eq:
        lhs, rhs = eq.split("=>")
        rev = False
    elif "=" in eq:
        lhs, rhs = eq.split("=")
        rev = True
    else:
        raise ChemkinError(f"{what}: no '=' in the equation")
    r, tb_l, co_l = _split_side(lhs, species_set, what)
    p, tb_r, co_r = _split_side(rhs, species_set, what)
    if tb_l != tb_r or co_l != co_r:
        raise ChemkinError(f"{what}: third body / fall-off collider differs between the two sides")
    if tb_l and co_l:
        raise ChemkinError(f"{what}: both +M and (+M)")
    return dict(eq=eq, what=what, reac=r, prod=p, rev=rev, A=A, beta=beta, E=E,
                third=tb_l, collider=co_l, low=None, high=None, troe=None, sri=None, eff={}, ford={}, dup=False,
                rev_arr=None, plog=[], cheb=[], tcheb=None, pcheb=None)


def read_kinetics(text: str):
    """Elements, species and reaction records of a CHEMKIN kinetics file."""
    elements, species, reactions, units = [], [], [], ("CAL/MOLE", "MOLES")
    for key, head, body in _blocks(text):
        if key == "ELEMENTS":
            elements += [w.upper() for w in head]
        elif key == "SPECIES":
            species += [w.upper() for w in head]
        elif key == "THERMO":
            raise ChemkinError("THERMO data inside the kinetics file is not supported; pass a thermodynamics file")
        else:
            e_unit, a_unit = "CAL/MOLE", "MOLES"
            for w in head:
                w = w.upper()
                if w in E_CONVERSION:
                    e_unit = w
                elif w == "MOLES":
                    a_unit = w
                elif w in ("MOLECULES", "EVOLTS"):
                    raise ChemkinError(f"REACTIONS unit {w} is not supported")
                else:
                    raise ChemkinError(f"unknown REACTIONS unit {w}")
            units = (e_unit, a_unit)
            sset = set(species)
            cur = None
            for ln in body:
                u = ln.upper()
                head_word = u.split("/")[0].strip()
                if "/" not in u and head_word in ("DUP", "DUPLICATE"):
                    if cur is None:
                        raise ChemkinError("DUPLICATE before any reaction")
                    cur["dup"] = True
                    continue
                if "/" in u and "=" not in u.split("/")[0]:
                    if cur is None:
                        raise ChemkinError(f"auxiliary line before any reaction: {ln!r}")
                    rest = u
                    for m in _AUX.finditer(u):
                        kw, vals = m.group(1).strip(), m.group(2).split()
                        if kw in ("DUP", "DUPLICATE"):
                            cur["dup"] = True
                        elif kw == "LOW":
                            if len(vals) != 3:
                                raise ChemkinError(f"{cur['what']}: LOW needs 3 values")
                            cur["low"] = [float(v) for v in vals]
                        elif kw in ("HIGH", "REV"):
                            key = "high" if kw == "HIGH" else "rev_arr"
                            if len(vals) != 3:
                                raise ChemkinError(f"{cur['what']}: {kw} needs 3 values")
                            if cur[key] is not None:
                                raise ChemkinError(f"{cur['what']}: the {kw} option is used more than once")
                            cur[key] = [float(v) for v in vals]
                        elif kw == "CHEB":
                            cur["cheb"] += [float(v) for v in vals]
                        elif kw in ("TCHEB", "PCHEB"):
                            if len(vals) != 2:
                                raise ChemkinError(f"{cur['what']}: {kw} needs 2 values")
                            cur[kw.lower()] = [float(v) for v in vals]
                        elif kw == "PLOG":
                            if len(vals) != 4:
                                raise ChemkinError(f"{cur['what']}: PLOG needs 4 values (P, A, beta, E)")
                            cur["plog"].append([float(v) for v in vals])
                        elif kw == "TROE":
                            if len(vals) not in (3, 4):
                                raise ChemkinError(f"{cur['what']}: TROE needs 3 or 4 values")
                            cur["troe"] = [float(v) for v in vals] + ([0.0] if len(vals) == 3 else [])
                        elif kw == "SRI":
                            # 3 parameters: d = 1, e = 0 (KineticsMap_CHEMKIN.hpp:749-767)
                            if len(vals) not in (3, 5):
                                raise ChemkinError(f"{cur['what']}: SRI needs 3 or 5 values")
                            cur["sri"] = [float(v) for v in vals] + ([1.0, 0.0] if len(vals) == 3 else [])
                        elif kw == "FORD":
                            # FORD /species order/ (ReactionPolicy_CHEMKIN.hpp:511-545)
                            if len(vals) != 2 or vals[0] not in sset:
                                raise ChemkinError(f"{cur['what']}: FORD needs a species and an order")
                            cur["ford"][vals[0]] = float(vals[1])
                        elif kw in _UNSUPPORTED:
                            raise ChemkinError(f"{cur['what']}: the {kw} option is not supported")
                        elif kw in sset:
                            if len(vals) != 1:
                                raise ChemkinError(f"{cur['what']}: efficiency of {kw} needs one value")
                            cur["eff"][kw] = float(vals[0])
                        else:
                            raise ChemkinError(f"{cur['what']}: unknown keyword or species {kw!r}")
                        rest = rest.replace(m.group(0), " ", 1)
                    leftover = rest.strip()
                    if leftover in ("DUP", "DUPLICATE"):
                        cur["dup"] = True
                    elif leftover:
                        raise ChemkinError(f"{cur['what']}: cannot read {leftover!r}")
                    continue
                cur = _parse_reaction_line(ln, sset, len(reactions))
                reactions.append(cur)
    if len(set(species)) != len(species):
        raise ChemkinError("species listed twice")
    return elements, species, reactions, units


# ----------------------------------------------------------------------------------
# tables
# ----------------------------------------------------------------------------------
def read_chemkin(kinetics: str, thermo: str, name: str = "chemkin") -> Mechanism:
    """Mechanism tables from the text of a CHEMKIN kinetics file and a thermodynamics file."""
    elements, species, reactions, (e_unit, _) = read_kinetics(kinetics)
    th = read_thermo(thermo)
    ns = len(species)
    idx = {s: k for k, s in enumerate(species)}
    missing = [s for s in species if s not in th]
    if missing:
        raise ChemkinError(f"no thermodynamic data for {missing}")
    el_names = list(dict.fromkeys(elements + [e for s in species for e in th[s][2]]))
    unknown = [e for e in el_names if e not in ATOMIC_WEIGHTS]
    if unknown:
        raise ChemkinError(f"no atomic weight for elements {unknown}")
    atoms = np.zeros((ns, len(el_names)))
    nasa = np.zeros((ns, 18))
    aw = np.array([ATOMIC_WEIGHTS[e] for e in el_names])
    for k, s in enumerate(species):
        c, (tl, thi, tm), comp = th[s]
        for e, n in comp.items():
            atoms[k, el_names.index(e)] = n
        nasa[k, :14] = c
        nasa[k, 14:17] = (tl, thi, tm)
        nasa[k, 17] = float(atoms[k] @ aw)

    econv = E_CONVERSION[e_unit]
    lnA, beta, EoR, flags, reac, prod = [], [], [], [], [], []
    tb_rxn, tb_sp, tb_eff = [], [], []
    fo_rxn, fo_type, fo_inf, fo_troe, fo_single, fo_sp, fo_eff = [], [], [], [], [], [], []
    fo_sri = []
    ord_rows = []
    rev_arr = np.zeros((len(reactions), 3))
    pl_rxn, pl_lev = [], []
    ch_rxn, ch_dims, ch_lim, ch_a = [], [], [], []
    for j, r in enumerate(reactions):
        what = r["what"]
        lam = dict(r["reac"])
        if r["ford"]:
            lam.update(r["ford"])
            ord_rows.append(sorted((idx[s], v) for s, v in lam.items()))
        else:
            ord_rows.append([])
        order = sum(lam.values())             # sumLambdaReactants_ (ReactionPolicy_CHEMKIN.hpp:1360-1366)
        cheb = bool(r["cheb"])
        falloff = r["collider"] is not None and not cheb
        cabr = r["high"] is not None
        if falloff and (r["low"] is None) == (r["high"] is None):
            raise ChemkinError(f"{what}: a (+M) reaction needs exactly one of LOW and HIGH")
        if not falloff and not cheb and any(r[k] is not None for k in ("low", "high", "troe", "sri")):
            raise ChemkinError(f"{what}: LOW/HIGH/TROE/SRI on a reaction without (+M)")
        if r["rev_arr"] is not None and (not r["rev"] or falloff):
            raise ChemkinError(f"{what}: REV needs a reversible reaction without (+M)")
        if r["plog"] and (falloff or r["third"]):
            raise ChemkinError(f"{what}: PLOG cannot be used on third-body or fall-off reactions")
        if (r["tcheb"] is not None or r["pcheb"] is not None) and not cheb:
            raise ChemkinError(f"{what}: TCHEB/PCHEB without CHEB")
        if cheb:
            if r["collider"] != "M":
                raise ChemkinError(f"{what}: CHEB can be used only with (+M)")
            if any(r[k] is not None for k in ("low", "high", "troe", "sri")) or r["plog"] or r["eff"]:
                raise ChemkinError(f"{what}: CHEB cannot be combined with LOW/HIGH/TROE/SRI/PLOG or efficiencies")
        if r["troe"] is not None and r["sri"] is not None:
            raise ChemkinError(f"{what}: both TROE and SRI")
        if r["sri"] is not None and (r["sri"][3] <= 0.0 or r["sri"][2] == 0.0):
            raise ChemkinError(f"{what}: SRI needs d > 0 and c != 0")
        if falloff:
            if cabr:
                # chemically activated: direct = the reaction line (order), k_inf = HIGH
                # (order - 1) (ReactionPolicy_CHEMKIN.hpp:1396-1404)
                (A0, b0, E0), (Ai, bi, Ei) = (r["A"], r["beta"], r["E"]), r["high"]
                c0, ci = 1e3 ** (1.0 - order), 1e3 ** (2.0 - order)
            else:
                # direct parameters: LOW (order + 1); k_inf: the Arrhenius line (order)
                (A0, b0, E0), (Ai, bi, Ei) = r["low"], (r["A"], r["beta"], r["E"])
                c0, ci = 1e3 ** (-order), 1e3 ** (1.0 - order)
            if Ai < 0.0 or A0 < 0.0:
                raise ChemkinError(f"{what}: negative pre-exponential factor on a fall-off reaction")
            lnA.append(np.log(A0 * c0))
            beta.append(b0)
            EoR.append(E0 * econv / R_J_KMOL)
            fo_rxn.append(j)
            fo_inf.append((np.log(Ai * ci), bi, Ei * econv / R_J_KMOL))
            fo_type.append((FALLOFF_TROE if r["troe"] is not None else
                            FALLOFF_SRI if r["sri"] is not None else FALLOFF_LINDEMANN) |
                           (FALLOFF_CHEMACT if cabr else 0))
            fo_troe.append(tuple(r["troe"]) if r["troe"] is not None else (0.0, 1.0, 1.0, 0.0))
            fo_sri.append(tuple(r["sri"]) if r["sri"] is not None else (0.0, 0.0, 1.0, 1.0, 0.0))
            if r["collider"] == "M":
                fo_single.append(-1)
                items = sorted((idx[s], e - 1.0) for s, e in r["eff"].items())
            else:
                if r["collider"] not in idx:
                    raise ChemkinError(f"{what}: unknown collider {r['collider']}")
                if r["eff"]:
                    raise ChemkinError(f"{what}: efficiencies given for a single-collider fall-off reaction")
                fo_single.append(idx[r["collider"]])
                items = []
            fo_sp.append([k for k, _ in items])
            fo_eff.append([e for _, e in items])
            fl = 0
        else:
            ordr = order + (1.0 if r["third"] else 0.0)
            A = r["A"]
            fl = RXN_NEGATIVE_A if A < 0.0 else 0
            # A = 0 is kept as ln(1e-100), like the preprocessor's writer
            # (preprocessing/PreProcessorKineticsPolicy_CHEMKIN.hpp:2142)
            lnA.append(np.log(abs(A) * 1e3 ** (1.0 - ordr)) if A != 0.0 else np.log(1e-100))
            beta.append(r["beta"])
            EoR.append(r["E"] * econv / R_J_KMOL)
            if r["third"]:
                tb_rxn.append(j)
                items = sorted((idx[s], e - 1.0) for s, e in r["eff"].items())
                tb_sp.append([k for k, _ in items])
                tb_eff.append([e for _, e in items])
            elif r["eff"]:
                raise ChemkinError(f"{what}: efficiencies on a reaction without M")
            if cheb:
                cv = r["cheb"]
                if len(cv) < 3 or int(cv[0]) * int(cv[1]) != len(cv) - 2 or cv[0] < 1 or cv[1] < 1:
                    raise ChemkinError(f"{what}: CHEB needs N M and N x M coefficients")
                if r["A"] < 0.0:
                    raise ChemkinError(f"{what}: negative pre-exponential factor on a Chebyshev reaction")
                Tl = r["tcheb"] or [300.0, 2500.0]
                Pl = r["pcheb"] or [0.001, 100.0]
                ch_rxn.append(j)
                ch_dims.append((int(cv[0]), int(cv[1])))
                ch_lim.append((Tl[0], Tl[1], Pl[0] * 101325.0, Pl[1] * 101325.0, 1e3 ** (1.0 - order)))
                ch_a.append(cv[2:])
            if r["rev_arr"] is not None:
                Ar, br, Er = r["rev_arr"]
                if Ar < 0.0:
                    raise ChemkinError(f"{what}: negative pre-exponential factor in REV")
                prod_order = sum(nu for _, nu in r["prod"]) + (1.0 if r["third"] else 0.0)
                rev_arr[j] = (np.log(Ar * 1e3 ** (1.0 - prod_order)) if Ar != 0.0 else np.log(1e-100), br,
                              Er * econv / R_J_KMOL)
                fl |= RXN_EXPLICIT_REV
            if r["plog"]:
                lev = []
                for P, Ap, bp, Ep in r["plog"]:
                    if Ap < 0.0:
                        raise ChemkinError(f"{what}: PLOG needs non-negative frequency factors")
                    lev.append((P * 101325.0, np.log(max(Ap * 1e3 ** (1.0 - order), 1e-32)), bp,
                                Ep * econv / R_J_KMOL))
                if any(b[0] <= a[0] for a, b in zip(lev[:-1], lev[1:])):
                    raise ChemkinError(f"{what}: PLOG pressures must increase")
                pl_rxn.append(j)
                pl_lev.append(lev)
        if r["rev"]:
            fl |= RXN_REVERSIBLE
        flags.append(fl)
        reac.append(sorted((idx[s], nu) for s, nu in r["reac"]))
        prod.append(sorted((idx[s], nu) for s, nu in r["prod"]))

    def csr(lists):
        ptr = np.zeros(len(lists) + 1, dtype=np.int32)
        ptr[1:] = np.cumsum([len(x) for x in lists])
        return ptr

    nfo = len(fo_rxn)
    return Mechanism(
        species=list(species), elements=atoms, nasa=nasa,
        lnA=np.array(lnA), beta=np.array(beta), E_over_R=np.array(EoR), rflags=np.array(flags),
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
        name=name, element_names=tuple(el_names), atomic_weights=aw,
        fo_sri=np.array(fo_sri, dtype=np.float64).reshape(nfo, 5),
        ord_ptr=csr(ord_rows), ord_sp=np.array([s for x in ord_rows for s, _ in x], dtype=np.int32),
        ord_val=np.array([v for x in ord_rows for _, v in x], dtype=np.float64),
        rev_arr=rev_arr, pl_rxn=np.array(pl_rxn, dtype=np.int32), pl_ptr=csr(pl_lev),
        pl_par=np.array([v for x in pl_lev for v in x], dtype=np.float64).reshape(-1, 4),
        ch_rxn=np.array(ch_rxn, dtype=np.int32), ch_dims=np.array(ch_dims, dtype=np.int32).reshape(-1, 2),
        ch_lim=np.array(ch_lim, dtype=np.float64).reshape(-1, 5), ch_ptr=csr(ch_a),
        ch_a=np.array([v for x in ch_a for v in x], dtype=np.float64),
    )


def load_chemkin(kinetics_path: str, thermo_path: str, name: str | None = None) -> Mechanism:
    """:func:`read_chemkin` on two files."""
    with open(kinetics_path) as f:
        kin = f.read()
    with open(thermo_path) as f:
        thermo = f.read()
    return read_chemkin(kin, thermo, name or os.path.splitext(os.path.basename(kinetics_path))[0])
