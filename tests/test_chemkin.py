"""CHEMKIN reader (laminarsmoke_amd/chemkin.py) on the reference's H2/CO, H2 and CH4
skeletal mechanisms.

The fixtures under tests/golden/ are the reference's own input files
(run/kinetic-mechanisms/POLIMI_H2CO_1412, POLIMI_H2_1412 with the thermodynamics file
cut down to the species used; POLIMI_CH4_SKELETAL_1412 as it is).  What pins the reader:

* structure read off the files by eye: species and reaction counts, which reactions
  are irreversible, third-body, fall-off (Troe / Lindemann), efficiencies;
* every reaction conserves every element (exactly, in the atom counts);
* NASA-7 data against published JANAF values (cp at 1000 K, formation enthalpy and
  entropy at 298.15 K) - a misplaced coefficient column fails these by far;
* rate constants against the file's numbers converted with physical unit factors
  (1 cm3/mol = 1e-3 m3/kmol, 1 cal = 4.18443 J as the reference defines it);
* the SRI fall-off rate of the CH4 skeletal mechanism against the textbook SRI form
  evaluated from the file's numbers;
* the one-step global mechanisms (explicit FORD orders): the orders, the A conversion
  with the order sum, and the rate k [CH4][O2];
* orders below 1 (fractional coefficients or FORD) take the reference's smoothed
  form: (c + 1e-9)^lambda well above 1.5e-5 kmol/m3, linear in c well below 9e-6;
* option and unit handling on small inline mechanisms, and a clear error for the
  options the kernels do not evaluate.
"""
import os

import numpy as np
import pytest

from laminarsmoke_amd.chemkin import ChemkinError, load_chemkin, read_chemkin, read_thermo
from laminarsmoke_amd.mechanism import (FALLOFF_CHEMACT, FALLOFF_LINDEMANN, FALLOFF_SRI, FALLOFF_TROE,
                                        RXN_EXPLICIT_REV, RXN_NEGATIVE_A,
                                        RXN_REVERSIBLE)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
THERMO = os.path.join(GOLDEN, "POLIMI_H2CO_1412_thermo.CKT")
R = 8.3144621                       # J/mol/K (math/PhysicalConstants.h:20)
CAL = 4.18443                       # J/cal (math/Conversions.h:112)


@pytest.fixture(scope="module")
def h2co():
    return load_chemkin(os.path.join(GOLDEN, "POLIMI_H2CO_1412.CKI"), THERMO)


@pytest.fixture(scope="module")
def h2():
    return load_chemkin(os.path.join(GOLDEN, "POLIMI_H2_1412.CKI"), THERMO)


@pytest.fixture(scope="module")
def ch4():
    return load_chemkin(os.path.join(GOLDEN, "POLIMI_CH4_SKELETAL_1412.CKI"),
                        os.path.join(GOLDEN, "POLIMI_CH4_SKELETAL_1412_thermo.CKT"))


def test_h2co_structure(h2co):
    m = h2co
    assert m.ns == 14 and m.nr == 33
    assert m.species[:4] == ["H2", "O2", "H2O", "H2O2"] and m.species[-3:] == ["HE", "AR", "N2"]
    irrev = np.where((m.rflags & RXN_REVERSIBLE) == 0)[0].tolist()
    assert irrev == [26, 30, 31, 32]                  # HCO+HO2=>H+OH+CO2 and the last three
    assert not np.any(m.rflags & RXN_NEGATIVE_A)
    assert m.tb_rxn.tolist() == [7, 8, 9, 13, 21]      # H2+M, O2+M, H+OH+M, O+OH+M, HCO+M
    assert m.fo_rxn.tolist() == [2, 12, 16]
    assert m.fo_type.tolist() == [FALLOFF_TROE, FALLOFF_TROE, FALLOFF_LINDEMANN]
    np.testing.assert_array_equal(m.fo_troe[0], [0.5, 1e-30, 1e30, 0.0])       # 3-parameter TROE
    np.testing.assert_array_equal(m.fo_troe[1], [0.7346, 94.0, 1756.0, 5182.0])
    # efficiencies of H+O2(+M): H2O 18, H2 2.5, N2 1.26, AR .8, HE .8, CO 1.2, CO2 2.4 (stored as eff-1)
    sp = [m.species[k] for k in m.fo_sp[m.fo_ptr[0]:m.fo_ptr[1]]]
    eff = dict(zip(sp, m.fo_eff_m1[m.fo_ptr[0]:m.fo_ptr[1]] + 1.0))
    assert eff == pytest.approx(dict(H2O=18.0, H2=2.5, N2=1.26, AR=0.8, HE=0.8, CO=1.2, CO2=2.4))
    # OH+OH=O+H2O: the repeated reactant is one entry with coefficient 2
    j = 6
    assert m.reac_nu[m.reac_ptr[j]:m.reac_ptr[j + 1]].tolist() == [2.0]
    # the DUPLICATE pair CO+OH=CO2+H is kept as two reactions
    assert m.lnA[17] != m.lnA[18] and m.reac_sp[m.reac_ptr[17]:m.reac_ptr[18]].tolist() == \
        m.reac_sp[m.reac_ptr[18]:m.reac_ptr[19]].tolist()


def test_h2_structure(h2):
    assert h2.ns == 9 and h2.nr == 20
    assert h2.element_names == ("H", "O", "N")


def test_ch4_structure(ch4):
    m = ch4
    assert m.ns == 24 and m.nr == 155 and len(m.tb_rxn) == 9 and len(m.fo_rxn) == 8
    assert m.species[:3] == ["CH4", "O2", "CO2"] and m.species[-1] == "N2"
    k = m.fo_type.tolist().index(FALLOFF_SRI)
    assert m.fo_type.tolist().count(FALLOFF_SRI) == 1 and m.fo_rxn[k] == 20      # H+CH3(+M)=CH4(+M)
    np.testing.assert_array_equal(m.fo_sri[k], [0.45, 797.0, 979.0, 1.0, 0.0])


@pytest.mark.parametrize("which", ["h2co", "h2", "ch4"])
def test_element_balance(request, which):
    m = request.getfixturevalue(which)
    bal = m.elements.T @ m.stoich_matrix()
    assert np.abs(bal).max() == 0.0


def test_molecular_weights(h2co):
    mw = dict(zip(h2co.species, h2co.mw))
    assert mw["H2O"] == pytest.approx(18.015, abs=2e-3)
    assert mw["CO2"] == pytest.approx(44.009, abs=2e-3)
    assert mw["AR"] == pytest.approx(39.948, abs=1e-5)       # single-precision table values
    assert mw["HE"] == pytest.approx(4.003, abs=1e-5)


# JANAF (NIST-JANAF Thermochemical Tables): cp(1000 K) J/mol/K, dHf(298.15) kJ/mol, S(298.15) J/mol/K
JANAF = {
    "H2": (30.203, 0.0, 130.680), "O2": (34.870, 0.0, 205.147), "N2": (32.697, 0.0, 191.609),
    "H2O": (41.268, -241.826, 188.834), "CO": (33.183, -110.527, 197.653),
    "CO2": (54.308, -393.522, 213.795), "AR": (20.786, 0.0, 154.845),
}


def test_thermo_against_janaf(h2co, oracle):
    for T in (298.15, 1000.0):
        cp, h, s = oracle.species_thermo(h2co, T)
        for name, (cp1000, hf, s298) in JANAF.items():
            k = h2co.species.index(name)
            if T == 1000.0:
                assert cp[k] * R == pytest.approx(cp1000, rel=5e-3), name
            else:
                assert h[k] * R * T / 1e3 == pytest.approx(hf, abs=0.5), name
                assert s[k] * R == pytest.approx(s298, rel=3e-3), name


def test_rate_constants(h2co, oracle):
    m = h2co
    T = 1300.0
    c = np.linspace(1e-3, 3e-3, m.ns)                 # kmol/m3
    k = oracle.kinetics(m, T, c)
    cgs2 = 1e-6 * 1e3                                  # cm3/mol/s -> m3/kmol/s
    # H+O2=OH+O  .9600E+15 -.200 16625.0
    assert k["kf"][0] == pytest.approx(0.96e15 * cgs2 * T ** -0.2 * np.exp(-16625.0 * CAL / (R * T)), rel=1e-12)
    # H2+M=H+H+M  .1115E+15 0 96081: second order counting M (cm3/mol/s)
    j = 7
    kf = 0.1115e15 * cgs2 * np.exp(-96081.0 * CAL / (R * T))
    assert k["kf"][j] == pytest.approx(kf, rel=1e-12)
    ck = dict(zip(m.species, c))
    M = c.sum() + 1.5 * ck["H2"] + 11.0 * ck["H2O"] - 0.5 * ck["AR"] + 0.9 * ck["CO"] + 2.8 * ck["CO2"]
    assert k["kmod"][j] == pytest.approx(kf * M, rel=1e-12)
    # CO+O(+M)=CO2(+M): k_inf .9640E+11 0 3800, LOW .2070E+27 -3.34 7610 (Lindemann)
    j = 16
    kinf = 0.964e11 * cgs2 * np.exp(-3800.0 * CAL / (R * T))
    k0 = 0.207e27 * 1e-6 * T ** -3.34 * np.exp(-7610.0 * CAL / (R * T))
    M = c.sum() + 11.0 * ck["H2O"] + 1.0 * ck["H2"] + 0.5 * ck["CO"] + 1.0 * ck["CO2"] - 0.5 * ck["AR"]
    Pr = k0 * M / kinf
    assert k["kf"][j] == pytest.approx(k0, rel=1e-12)
    assert k["kmod"][j] == pytest.approx(kinf * Pr / (1.0 + Pr), rel=1e-12)


def test_sri_rate(ch4, oracle):
    """H+CH3(+M)=CH4(+M) 1.2E15 -0.4 0 / LOW 6.4E23 -1.8 0 / SRI 0.45 797 979 1 0 /
    efficiencies H2 2, CO 2, CO2 3, H2O 5."""
    m = ch4
    j = 20
    ck = {s: 1e-3 * (1 + 0.1 * i) for i, s in enumerate(m.species)}
    c = np.array([ck[s] for s in m.species])
    for T in (800.0, 1500.0, 2300.0):
        k = oracle.kinetics(m, T, c)
        kinf = 1.2e15 * 1e-3 * T ** -0.4
        k0 = 6.4e23 * 1e-6 * T ** -1.8
        M = c.sum() + ck["H2"] + ck["CO"] + 2.0 * ck["CO2"] + 4.0 * ck["H2O"]
        Pr = k0 * M / kinf
        X = 1.0 / (1.0 + np.log10(Pr) ** 2)
        F = (0.45 * np.exp(-797.0 / T) + np.exp(-T / 979.0)) ** X
        assert k["kf"][j] == pytest.approx(k0, rel=1e-12)
        assert k["kmod"][j] == pytest.approx(kinf * Pr / (1.0 + Pr) * F, rel=1e-12)


def test_sri_five_parameter_form():
    m = _small("H+O2(+M)=HO2(+M) 1.0E12 0.5 0.0\n LOW/ 1.0E18 -1.0 0.0/\n SRI/ 0.5 100 1000 2.0 0.3/")
    assert m.fo_type.tolist() == [FALLOFF_SRI]
    np.testing.assert_array_equal(m.fo_sri[0], [0.5, 100.0, 1000.0, 2.0, 0.3])


def test_thermo_reader_fields():
    th = read_thermo(open(THERMO).read())
    coef, (tl, thi, tm), comp = th["AR"]
    assert (tl, thi, tm) == (300.0, 3500.0, 1490.0) and comp == {"AR": 1.0}
    assert coef[0] == 2.5 and coef[5] == -745.375 and coef[13] == 4.366
    assert th["H2O2"][2] == {"H": 2.0, "O": 2.0}


def _small(reactions, units=""):
    kin = f"ELEMENTS H O N END\nSPECIES H2 O2 H2O H O OH HO2 N2 END\nREACTIONS {units}\n{reactions}\nEND\n"
    return read_chemkin(kin, open(THERMO).read())


def test_energy_units():
    base = _small("H+O2<=>OH+O 1.0E13 0.0 16.625", "KCAL/MOLE").E_over_R[0]
    assert base == pytest.approx(16625.0 * CAL / R, rel=1e-14)
    assert _small("H+O2=OH+O 1.0E13 0.0 8000.0", "KELVINS").E_over_R[0] == pytest.approx(8000.0, rel=1e-14)
    assert _small("H+O2=OH+O 1.0E13 0.0 69.565", "KJOULES/MOLE").E_over_R[0] == \
        pytest.approx(69565.0 / R, rel=1e-14)


def test_coefficients_negative_A_and_collider():
    m = _small("2OH=>O+H2O 1.0E12 0.0 0.0\nH+OH=H2O -1.0E12 0.0 0.0\n"
               "H+O2(+N2)=HO2(+N2) 1.0E12 0.5 0.0\n LOW/ 1.0E18 -1.0 0.0/")
    assert m.reac_nu[m.reac_ptr[0]:m.reac_ptr[1]].tolist() == [2.0]
    assert m.rflags[0] == 0 and m.rflags[1] == RXN_NEGATIVE_A | RXN_REVERSIBLE
    assert m.lnA[1] == pytest.approx(np.log(1e12 * 1e-3))
    assert m.fo_single.tolist() == [m.species.index("N2")] and m.fo_type.tolist() == [FALLOFF_LINDEMANN]
    assert m.lnA[2] == pytest.approx(np.log(1e18 * 1e-6)) and m.fo_inf[0, 0] == pytest.approx(np.log(1e12 * 1e-3))


def test_rev_high_plog_units():
    """REV converts with the product order (+1 with M), HIGH with order - 1 (the
    reaction line of a chemically activated reaction is k0 with the reaction
    order), PLOG levels like the reaction line, P in atm -> Pa
    (ReactionPolicy_CHEMKIN.hpp:1359-1376, 1396-1404, 1459-1478)."""
    m = _small("H+O2=OH+O 1.0E13 0.5 1000.0\n REV/ 2.0E12 -0.3 500.0/\n"
               "H+OH+M=H2O+M 1.0E22 -2.0 0.0\n REV/ 3.0E23 -1.5 120000.0/\n"
               "H+O2(+M)=HO2(+M) 5.0E12 0.2 300.0\n HIGH/ 4.0E14 0.1 900.0/\n TROE/ 0.5 100 2000 /\n"
               "H2+O=H+OH 1.0E99 0.0 0.0\n PLOG/ 0.1 1.0E12 0.5 2000.0/\n PLOG/ 10.0 4.0E13 0.0 3000.0/\n"
               "H2+O2=HO2+H 1.0E10 0.0 0.0\n PLOG/ 1.0 0.0 0.0 0.0/")
    assert m.rflags[0] == RXN_REVERSIBLE | RXN_EXPLICIT_REV
    np.testing.assert_allclose(m.rev_arr[0], [np.log(2e12 * 1e-3), -0.3, 500.0 * CAL / R], rtol=1e-14)
    # third body, 1 product + M: order 2
    np.testing.assert_allclose(m.rev_arr[1], [np.log(3e23 * 1e-3), -1.5, 120000.0 * CAL / R], rtol=1e-14)
    assert m.fo_type.tolist() == [FALLOFF_TROE | FALLOFF_CHEMACT]
    assert m.lnA[2] == pytest.approx(np.log(5e12 * 1e-3), rel=1e-14)           # k0: order 2
    assert m.fo_inf[0, 0] == pytest.approx(np.log(4e14), rel=1e-14)             # k_inf: order 1
    assert m.pl_rxn.tolist() == [3, 4] and m.pl_ptr.tolist() == [0, 2, 3]
    np.testing.assert_allclose(m.pl_par[:2], [[0.1 * 101325.0, np.log(1e12 * 1e-3), 0.5, 2000.0 * CAL / R],
                                              [10.0 * 101325.0, np.log(4e13 * 1e-3), 0.0, 3000.0 * CAL / R]],
                               rtol=1e-14)
    assert m.pl_par[2, 1] == pytest.approx(np.log(1e-32))                       # A = 0 floored


@pytest.mark.parametrize("text,msg", [
    ("H+O2(+M)=HO2(+M) 1.0E12 0.5 0.0\n LOW/ 1.0E18 -1.0 0.0/\n SRI/ 0.5 100 1000 0.0 0.0/", "SRI needs d > 0"),
    ("H+O2(+M)=HO2(+M) 1.0E12 0.5 0.0\n HIGH/ 1.0E18 -1.0 0.0/\n LOW/ 1.0E18 -1.0 0.0/", "one of LOW and HIGH"),
    ("H+O2=HO2 1.0E12 0.5 0.0\n HIGH/ 1.0E18 -1.0 0.0/", "without \\(\\+M\\)"),
    ("H2+O2=>H2O+O 1.0E12 0.0 0.0\n RORD /H2 1.5/", "RORD"),
    ("H+O2=>OH+O 1.0E12 0.0 0.0\n REV/ 1.0E11 0.0 0.0/", "REV needs a reversible"),
    ("H+O2=OH+O 1.0E12 0.0 0.0\n REV/ 1.0E11 0.0 0.0/\n REV/ 1.0E11 0.0 0.0/", "more than once"),
    ("H+O2(+M)=HO2(+M) 1.0E12 0.5 0.0\n LOW/ 1.0E18 -1.0 0.0/\n REV/ 1.0E11 0.0 0.0/", "REV needs"),
    ("H+O2+M=HO2+M 1.0E12 0.0 0.0\n PLOG/ 1.0 1.0E12 0.0 0.0/", "PLOG cannot"),
    ("H+O2=OH+O 1.0E12 0.0 0.0\n PLOG/ 10.0 1.0E12 0.0 0.0/\n PLOG/ 1.0 1.0E12 0.0 0.0/", "must increase"),
    ("H+O2=OH+O 1.0E12 0.0 0.0\n PLOG/ 1.0 -1.0E12 0.0 0.0/", "non-negative"),
    ("H+O2=OH+O 1.0E12 0.0 0.0\n CHEB/ 1 1 3.0/", "CHEB can be used only with"),
    ("H+O2(+M)=HO2(+M) 1.0E12 0.0 0.0\n CHEB/ 2 2 3.0 1.0 0.5/", "N x M coefficients"),
    ("H+O2(+M)=HO2(+M) 1.0E12 0.0 0.0\n LOW/ 1.0E18 -1.0 0.0/\n CHEB/ 1 1 3.0/", "CHEB cannot"),
    ("H+O2=OH+O 1.0E12 0.0 0.0\n TCHEB/ 300 2500/", "without CHEB"),
    ("H+O2=OH+XX 1.0E12 0.0 0.0", "unknown species"),
    ("H+O2(+M)=HO2(+M) 1.0E12 0.5 0.0", "one of LOW and HIGH"),
])
def test_unsupported_options_raise(text, msg):
    with pytest.raises(ChemkinError, match=msg):
        _small(text)


def test_integration_conserves_elements_and_enthalpy(h2co, oracle):
    """Stoichiometric H2/CO/air at 1200 K ignites within 1 ms (oracle); elements and enthalpy conserved."""
    m = h2co
    x = np.zeros(m.ns)
    sp = m.species.index
    x[sp("H2")], x[sp("CO")], x[sp("O2")], x[sp("N2")] = 0.5, 0.5, 0.5, 0.5 * 3.76
    y0 = x * m.mw / (x @ m.mw)
    y = np.concatenate([y0, [1200.0]])[:, None]
    yf, st, q, cnt = oracle.chemistry_step(m, y, 1e-3, rtol=1e-8, atol=1e-14)
    assert st[0] > 0
    assert yf[-1, 0] > 2000.0
    em = m.element_mass_matrix()
    np.testing.assert_allclose(em @ yf[:m.ns, 0], em @ y0, atol=1e-12)
    cp0, h0, _ = oracle.species_thermo(m, 1200.0)
    cpf, hf, _ = oracle.species_thermo(m, yf[-1, 0])
    H0 = (y0 / m.mw) @ h0 * R * 1200.0
    Hf = (yf[:m.ns, 0] / m.mw) @ hf * R * yf[-1, 0]
    assert abs(Hf - H0) <= 1e-6 * abs(cp0 @ (y0 / m.mw)) * R * 1200.0


@pytest.mark.parametrize("name,fuel,a", [("GLOBAL_CH4_1step", "CH4", 1.0e15), ("GLOBAL_H2_1step", "H2", 1.0e15)])
def test_global_mechanisms(oracle, name, fuel, a):
    m = load_chemkin(os.path.join(GOLDEN, name + ".CKI"), os.path.join(GOLDEN, name + "_thermo.CKT"))
    assert m.nr == 1 and m.rflags[0] == 0
    orders = dict(zip([m.species[k] for k in m.ord_sp], m.ord_val))
    assert orders == {fuel: 1.0, "O2": 1.0}
    assert m.lnA[0] == pytest.approx(np.log(a * 1e-3), rel=1e-15)       # orders sum to 2: cm3/mol/s
    T = 1500.0
    c = np.linspace(2e-3, 4e-3, m.ns)
    k = oracle.kinetics(m, T, c)
    kf = a * 1e-3 * np.exp(-20000.0 * CAL / (R * T))
    cf, co = c[m.species.index(fuel)], c[m.species.index("O2")]
    assert k["r"][0] == pytest.approx(kf * cf * co, rel=1e-12)


def test_orders_below_one_are_smoothed(oracle):
    """A + 0.5 B => C: the B factor is (c + 1e-9)^0.5 at c = 1e-2 and linear in c at 1e-12."""
    m = _small("H2+0.5O2=>H2O 1.0E10 0.0 0.0")
    sp = m.species.index
    T = 1000.0
    c = np.full(m.ns, 1e-2)
    r_hi = oracle.kinetics(m, T, c)["r"][0]
    kf = 1.0e10 * 10 ** (-3 * 0.5) * 1.0      # order 1.5: (cm3/mol)^0.5 -> (m3/kmol)^0.5
    assert r_hi == pytest.approx(kf * 1e-2 * (1e-2 + 1e-9) ** 0.5, rel=1e-12)
    assert abs(r_hi / (kf * 1e-2 * np.sqrt(1e-2)) - 1.0) > 1e-8
    c[sp("O2")] = 1e-12
    r1 = oracle.kinetics(m, T, c)["r"][0]
    c[sp("O2")] = 2e-12
    r2 = oracle.kinetics(m, T, c)["r"][0]
    assert r2 / r1 == pytest.approx(2.0, rel=1e-6)
