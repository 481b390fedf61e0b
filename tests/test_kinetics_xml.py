"""kinetics.xml reader and writer (laminarsmoke_amd/kinetics_xml.py).

* A hand-written kinetics.xml in the layout the reference reads
  (KineticsMap_CHEMKIN::ImportCoefficientsFromXMLFile, StoichiometricMap::ReadFromASCIIFile,
  ThermodynamicsMap_CHEMKIN::ImportCoefficientsFromXMLFile): 1-based indices, raw
  third-body efficiencies (the maps keep eff - 1), 3-parameter Troe (T** = 0) and
  SRI (d = 1, e = 0) defaults, a single-collider fall-off, a negative pre-exponential
  factor, the stoichiometric classes 1, 2, 0.5 and "other".  Every table entry is
  checked against the values written into the file.
* Tables written and read back are identical, bit for bit, for the reference's CHEMKIN
  mechanisms and the synthetic workload mechanisms.
* Explicit forward orders (FORD of the one-step global mechanisms) travel in the
  lambda lists and come back as the same order tables.
* Forms the kernels do not evaluate are refused.
No kinetics.xml file ships with the reference (its mechanisms are CHEMKIN input), so
the format is pinned by the reader code it follows, not by a reference file.
"""
import os

import numpy as np
import pytest

from laminarsmoke_amd.chemkin import load_chemkin
from laminarsmoke_amd.kinetics_xml import KineticsXmlError, read_kinetics_xml, write_kinetics_xml
from laminarsmoke_amd.mechanism import (FALLOFF_SRI, FALLOFF_TROE, RXN_NEGATIVE_A, RXN_REVERSIBLE,
                                        workload_mechanism)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

NASA = " ".join(str(float(i)) for i in range(18))

# species A, B, C; reactions:
#  1  A + B = 2 C         (reversible, third body, efficiencies A 2.5, C 0.5)
#  2  2 A => B            (irreversible, negative A)
#  3  0.5 B + A = C (+B)  (reversible, Troe with 3 parameters, collider B)
#  4  A = 1.7 C (+M)      (reversible, SRI with 3 parameters, efficiency C 3.0)
HAND = f"""<?xml version="1.0" encoding="utf-8"?>
<opensmoke version="0.1a">
<NumberOfElements> 1 </NumberOfElements>
<NamesOfElements> H </NamesOfElements>
<AtomicComposition> 2 4 2 </AtomicComposition>
<NumberOfSpecies> 3 </NumberOfSpecies>
<NamesOfSpecies> A B C </NamesOfSpecies>
<Thermodynamics type="NASA"><NASA-coefficients>
{NASA} {NASA} {NASA}
</NASA-coefficients></Thermodynamics>
<Kinetics type="OpenSMOKE" version="04-22-2013">
<NumberOfReactions> 4 </NumberOfReactions>
<Irreversible> 1 2 </Irreversible>
<Reversible> 3 1 3 4 </Reversible>
<Reversible-Thermodynamics> 3 1 3 4 </Reversible-Thermodynamics>
<Reversible-Explicit> 0 </Reversible-Explicit>
<ThreeBody> 1 1 </ThreeBody>
<FallOff> 2 3 4 <Lindemann> 0 </Lindemann> <Troe> 1 3 </Troe> <SRI> 1 4 </SRI> </FallOff>
<CABR> 0 <Lindemann> 0 </Lindemann> <Troe> 0 </Troe> <SRI> 0 </SRI> </CABR>
<KineticParameters>
<Direct>
<lnA> 4 1.5 2.5 3.5 4.5 </lnA>
<negative-lnA> 1 2 </negative-lnA>
<Beta> 4 0.1 0.2 0.3 0.4 </Beta>
<E_over_R> 4 100 200 300 400 </E_over_R>
</Direct>
<ThreeBody><reaction i="1"> 2 1 3 2 2.5 0.5 </reaction></ThreeBody>
<FallOff>
<lnA> 2 10.5 11.5 </lnA> <Beta> 2 -0.5 -0.6 </Beta> <E_over_R> 2 50 60 </E_over_R>
<Parameters>
<reaction i="3"> troe 3 0.6 100 1000 species 2 </reaction>
<reaction i="4"> sri 3 0.45 797 979 thirdbody 1 3 1 3.0 </reaction>
</Parameters>
</FallOff>
</KineticParameters>
<Stoichiometry type="OpenSMOKE" version="04-22-2013">
3 3 1 0
3 1 0 0
3 0 0 0
3 0 1 0
3 0 0 0
3 0 1 1
3 0 0 1
3 0 0 0
3 0 0 0
3 0 0 1
3 0 0 1
3 0 0 1
3 0 0 0
3 0 0 0
3 0 0 1
4 1 3 4 1
1 2
0
1 3
0
0
2 2 3
1 1
0
0
1 4
1 1.7
1 3
1 1
0
0
1 4
1 1.7
4 0 -1 -0.5 0.7
0
</Stoichiometry>
</Kinetics>
</opensmoke>
"""


def test_hand_written_file():
    m = read_kinetics_xml(HAND)
    assert m.species == ["A", "B", "C"] and m.nr == 4
    np.testing.assert_array_equal(m.nasa, np.tile(np.arange(18.0), (3, 1)))
    np.testing.assert_array_equal(m.lnA, [1.5, 2.5, 3.5, 4.5])
    np.testing.assert_array_equal(m.beta, [0.1, 0.2, 0.3, 0.4])
    np.testing.assert_array_equal(m.E_over_R, [100, 200, 300, 400])
    assert m.rflags.tolist() == [RXN_REVERSIBLE, RXN_NEGATIVE_A, RXN_REVERSIBLE, RXN_REVERSIBLE]
    side = lambda ptr, sp, nu, j: list(zip(sp[ptr[j]:ptr[j + 1]].tolist(), nu[ptr[j]:ptr[j + 1]].tolist()))
    assert side(m.reac_ptr, m.reac_sp, m.reac_nu, 0) == [(0, 1.0), (1, 1.0)]
    assert side(m.prod_ptr, m.prod_sp, m.prod_nu, 0) == [(2, 2.0)]
    assert side(m.reac_ptr, m.reac_sp, m.reac_nu, 1) == [(0, 2.0)]
    assert side(m.prod_ptr, m.prod_sp, m.prod_nu, 1) == [(1, 1.0)]
    assert side(m.reac_ptr, m.reac_sp, m.reac_nu, 2) == [(0, 1.0), (1, 0.5)]
    assert side(m.prod_ptr, m.prod_sp, m.prod_nu, 2) == [(2, 1.0)]
    assert side(m.reac_ptr, m.reac_sp, m.reac_nu, 3) == [(0, 1.0)]
    assert side(m.prod_ptr, m.prod_sp, m.prod_nu, 3) == [(2, 1.7)]
    assert m.tb_rxn.tolist() == [0] and m.tb_sp.tolist() == [0, 2]
    np.testing.assert_array_equal(m.tb_eff_m1, [1.5, -0.5])
    assert m.fo_rxn.tolist() == [2, 3] and m.fo_type.tolist() == [FALLOFF_TROE, FALLOFF_SRI]
    np.testing.assert_array_equal(m.fo_inf, [[10.5, -0.5, 50.0], [11.5, -0.6, 60.0]])
    np.testing.assert_array_equal(m.fo_troe[0], [0.6, 100.0, 1000.0, 0.0])
    np.testing.assert_array_equal(m.fo_sri[1], [0.45, 797.0, 979.0, 1.0, 0.0])
    assert m.fo_single.tolist() == [1, -1] and m.fo_sp.tolist() == [2]
    np.testing.assert_array_equal(m.fo_eff_m1, [2.0])
    np.testing.assert_array_equal(m.elements, [[2.0], [4.0], [2.0]])


def _same(a, b):
    for f in ("nasa", "lnA", "beta", "E_over_R", "rflags", "reac_ptr", "reac_sp", "reac_nu", "prod_ptr", "prod_sp",
              "prod_nu", "tb_rxn", "tb_ptr", "tb_sp", "tb_eff_m1", "fo_rxn", "fo_type", "fo_inf", "fo_troe",
              "fo_single", "fo_ptr", "fo_sp", "fo_eff_m1", "fo_sri", "ord_ptr", "ord_sp", "ord_val", "elements",
              "rev_arr", "pl_rxn", "pl_ptr", "ch_rxn", "ch_dims", "ch_ptr", "ch_a"):
        x, y = getattr(a, f), getattr(b, f)
        assert x.shape == y.shape and np.array_equal(x, y), f
    # PLOG levels travel as raw CHEMKIN values (P in atm, A, E) with unit factors
    assert a.pl_par.shape == b.pl_par.shape
    np.testing.assert_allclose(a.pl_par, b.pl_par, rtol=1e-14, atol=0)
    assert a.ch_lim.shape == b.ch_lim.shape                       # Chebyshev pressures travel in atm
    np.testing.assert_allclose(a.ch_lim, b.ch_lim, rtol=1e-15, atol=0)
    assert a.species == b.species


@pytest.mark.parametrize("name", ["POLIMI_H2CO_1412", "POLIMI_H2_1412", "POLIMI_CH4_SKELETAL_1412", "GLOBAL_CH4_1step",
                                  "GLOBAL_H2_1step", 14, 53, 160])
def test_round_trip(name):
    if isinstance(name, int):
        m = workload_mechanism(name)
    else:
        thermo = os.path.join(GOLDEN, name + "_thermo.CKT")
        if not os.path.exists(thermo):
            thermo = os.path.join(GOLDEN, "POLIMI_H2CO_1412_thermo.CKT")
        m = load_chemkin(os.path.join(GOLDEN, name + ".CKI"), thermo)
    _same(m, read_kinetics_xml(write_kinetics_xml(m)))


REV_CABR_PLOG = """ELEMENTS H O N END
SPECIES H2 O2 H2O H O OH HO2 N2 H2O2 END
REACTIONS
H+O2=OH+O 1.0E13 0.5 1000.0
 REV/ 2.0E12 -0.3 500.0/
H+O2(+M)=HO2(+M) 5.0E12 0.2 300.0
 HIGH/ 4.0E14 0.1 900.0/
 TROE/ 0.5 100 2000 /
 H2O/12.0/ N2/0.8/
H2+O=H+OH 1.0E12 0.0 0.0
 PLOG/ 0.1 1.0E12 0.5 2000.0/
 PLOG/ 10.0 4.0E13 0.0 3000.0/
H+OH+M=H2O+M 1.0E22 -2.0 0.0
 REV/ 3.0E23 -1.5 120000.0/
H+HO2(+N2)=H2+O2(+N2) 1.0E12 0.0 0.0
 LOW/ 1.0E16 0.0 0.0/
H+HO2(+N2)=2OH(+N2) 1.0E11 0.0 0.0
 HIGH/ 1.0E9 0.0 0.0/
 SRI/ 0.5 200 900 /
H2O2(+M)=2OH(+M) 1.0 0.0 0.0
 TCHEB/ 290.0 3000.0 / PCHEB/ 0.01 98.7 /
 CHEB/ 3 2 8.0 0.5 -1.2 0.1 0.05 -0.02 /
END
"""


def test_round_trip_rev_cabr_plog(tmp_path):
    """Explicit reverse, chemically activated (between two fall-off entries), PLOG and
    Chebyshev reactions come back as the same tables; the XML holds the reference's lists."""
    from laminarsmoke_amd.chemkin import read_chemkin
    from laminarsmoke_amd.mechanism import FALLOFF_CHEMACT, RXN_EXPLICIT_REV
    m = read_chemkin(REV_CABR_PLOG, open(os.path.join(GOLDEN, "POLIMI_H2CO_1412_thermo.CKT")).read())
    assert m.fo_type.tolist() == [FALLOFF_TROE | FALLOFF_CHEMACT, 0, FALLOFF_SRI | FALLOFF_CHEMACT]
    assert (m.rflags & RXN_EXPLICIT_REV).nonzero()[0].tolist() == [0, 3]
    text = write_kinetics_xml(m)
    for tag, body in (("Reversible-Explicit", "2\n1 4"), ("Reversible-Thermodynamics", "5\n2 3 5 6 7"),
                      ("PressureLog", "1\n3"), ("Chebyshev", "1\n7")):
        assert f"<{tag}>\n{body}\n</{tag}>" in text, tag
    assert "<CABR>\n2\n2 6\n" in text and "<FallOff>\n1\n5\n" in text
    assert m.ch_dims.tolist() == [[3, 2]] and m.ch_lim[0, 4] == 1.0        # unimolecular: conv 1e3^0
    np.testing.assert_array_equal(m.ch_a, [8.0, 0.5, -1.2, 0.1, 0.05, -0.02])
    _same(m, read_kinetics_xml(text))


@pytest.mark.parametrize("old,new,msg", [
    ("<Reversible-Explicit> 0 </Reversible-Explicit>", "<Reversible-Explicit> 1 2 </Reversible-Explicit>",
     "Reversible-Explicit"),
    ("<Reversible-Explicit> 0 </Reversible-Explicit>", "<Reversible-Explicit> 1 3 </Reversible-Explicit>",
     "<Reverse> not found"),
    ("<Reversible-Explicit> 0 </Reversible-Explicit>",
     "<Reversible-Explicit> 0 </Reversible-Explicit>\n<Chebyshev> 1 2 </Chebyshev>", "Chebyshev"),
    ("0.7\n0\n</Stoichiometry>", "0.7\n1\n</Stoichiometry>", "unexpected end"),   # flag without lambda lists
    ('version="04-22-2013">\n<NumberOfReactions>', 'version="01-01-2000">\n<NumberOfReactions>', "not supported"),
])
def test_refused(old, new, msg):
    assert HAND.count(old) == 1
    with pytest.raises(KineticsXmlError, match=msg):
        read_kinetics_xml(HAND.replace(old, new))
