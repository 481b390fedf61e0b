"""Pins of the oracle's thermodynamics, kinetics and reactor RHS (CPU only).

Every check compares the C oracle with something it does not share code with:
closed forms, thermodynamic identities, limits, hand-computed values, or the
independent dense formulation in tests/dense_reference.py.
"""
import numpy as np
import pytest

import dense_reference as dref
from laminarsmoke_amd.mechanism import (constant_cp_nasa, custom_mechanism, synthetic_mechanism,
                                        workload_mechanism)

R = 8314.4621


@pytest.mark.parametrize("ns", [14, 53, 160])
def test_nasa_identities(oracle, ns):
    """dh/dT = cp and ds/dT = cp/T (finite differences), continuity at T_mid."""
    m = workload_mechanism(ns)
    for T in (450.0, 900.0, 1500.0, 2400.0, 3400.0):
        d = 1e-3
        cp, h, s = oracle.species_thermo(m, T)
        _, hp, sp = oracle.species_thermo(m, T + d)
        _, hm, sm = oracle.species_thermo(m, T - d)
        dhdT = ((hp * (T + d)) - (hm * (T - d))) / (2 * d)       # d(h/R)/dT
        np.testing.assert_allclose(dhdT, cp, rtol=1e-7, atol=1e-7)
        np.testing.assert_allclose((sp - sm) / (2 * d), cp / T, rtol=1e-6, atol=1e-9)
    # continuity of cp, h, s across T_mid = 1000 K
    lo = oracle.species_thermo(m, 1000.0)
    hi = oracle.species_thermo(m, np.nextafter(1000.0, 2000.0))
    for a, b in zip(lo, hi):
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-9)


def test_nasa_matches_dense_formulation(oracle):
    m = workload_mechanism(53)
    for T in (300.0, 999.0, 1000.0, 1001.0, 2500.0, 5000.0):
        got = oracle.species_thermo(m, T)
        ref = dref.nasa_props(m, T)
        for a, b in zip(got, ref):
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-9)


def test_nasa_branch_is_strict(oracle):
    """High-T set iff T > T_mid (ThermodynamicsMap_CHEMKIN.hpp:449): at T_mid the low set is used."""
    row = constant_cp_nasa(3.5, 0.0, 20.0, 28.0)
    row[0] = 7.0   # high-T cp/R differs
    m = custom_mechanism(["A"], [row], [dict(reac=[(0, 1)], prod=[(0, 1)], lnA=0.0, reversible=False)])
    assert oracle.species_thermo(m, 1000.0)[0][0] == pytest.approx(3.5)
    assert oracle.species_thermo(m, 1000.0001)[0][0] == pytest.approx(7.0)


def test_constant_cp_closed_form(oracle):
    """cp constant => h = h0 + cp (T - T0), s = s0 + cp ln(T/T0)."""
    row = constant_cp_nasa(4.25, -1234.0, 25.5, 30.0)
    m = custom_mechanism(["A"], [row], [dict(reac=[(0, 1)], prod=[(0, 1)], lnA=0.0, reversible=False)])
    for T in (350.0, 1700.0):
        cp, h, s = oracle.species_thermo(m, T)
        assert cp[0] == pytest.approx(4.25, rel=1e-14)
        assert h[0] * T == pytest.approx(-1234.0 + 4.25 * (T - 298.15), rel=1e-12)
        assert s[0] == pytest.approx(25.5 + 4.25 * np.log(T / 298.15), rel=1e-12)


def _two_species_mech(**rxn):
    rows = [constant_cp_nasa(3.5, 0.0, 20.0, 20.0), constant_cp_nasa(4.0, -3000.0, 22.0, 20.0),
            constant_cp_nasa(4.5, 1500.0, 24.0, 20.0)]
    return custom_mechanism(["A", "B", "C"], rows, [rxn])


def test_arrhenius_and_negative_A(oracle):
    lnA, beta, EoR = np.log(3.7e7), -0.83, 5432.1
    m = _two_species_mech(reac=[(0, 1)], prod=[(1, 1)], lnA=lnA, beta=beta, E_over_R=EoR, reversible=False)
    c = np.array([0.3, 0.2, 0.1])
    for T in (500.0, 1800.0):
        k = oracle.kinetics(m, T, c)
        expect = 3.7e7 * T ** beta * np.exp(-EoR / T)
        assert k["kf"][0] == pytest.approx(expect, rel=1e-13)
        assert k["r"][0] == pytest.approx(expect * 0.3, rel=1e-13)
        np.testing.assert_allclose(k["R"], [-expect * 0.3, expect * 0.3, 0.0], rtol=1e-13)
    mneg = _two_species_mech(reac=[(0, 1)], prod=[(1, 1)], lnA=lnA, reversible=False, negative_A=True)
    assert oracle.kinetics(mneg, 900.0, c)["r"][0] == pytest.approx(-3.7e7 * 0.3, rel=1e-13)


def test_reaction_orders(oracle):
    """2A + C -> B: rate k c_A^2 c_C.  An order below 1 makes the side non-elementary
    (StoichiometricMap.hpp:256-300): its factor is (c + 1e-9)^0.5 at c = 0.7, where the
    smoothing weight is 1 (StoichiometricMap.hpp:617-654), not sqrt(c)."""
    m = _two_species_mech(reac=[(0, 2), (2, 1)], prod=[(1, 1)], lnA=np.log(11.0), reversible=False)
    c = np.array([0.7, 0.1, 0.3])
    assert oracle.kinetics(m, 1000.0, c)["r"][0] == pytest.approx(11.0 * 0.49 * 0.3, rel=1e-14)
    m = _two_species_mech(reac=[(0, 0.5)], prod=[(1, 0.5)], lnA=np.log(2.0), reversible=False)
    assert oracle.kinetics(m, 1000.0, c)["r"][0] == pytest.approx(2.0 * np.sqrt(0.7 + 1e-9), rel=1e-14)


@pytest.mark.parametrize("T", [700.0, 1600.0, 2900.0])
def test_detailed_balance(oracle, T):
    """At the composition where c_B c_C / c_A^2 equals Kc (from Gibbs energies), r = 0."""
    m = _two_species_mech(reac=[(0, 2)], prod=[(1, 1), (2, 1)], lnA=np.log(5e9), beta=0.3, E_over_R=1000.0)
    cp, h, s = dref.nasa_props(m, T)
    g = h - s
    Kc = np.exp(-(g[1] + g[2] - 2 * g[0]))            # dn = 0
    cA, cB = 0.02, 0.05
    cC = Kc * cA ** 2 / cB
    k = oracle.kinetics(m, T, np.array([cA, cB, cC]))
    fwd = k["kf"][0] * cA ** 2
    assert abs(k["r"][0]) < 1e-12 * fwd
    # dn != 0: A + B <=> C, Kc = Kp (Patm/RT)^-1
    m2 = _two_species_mech(reac=[(0, 1), (1, 1)], prod=[(2, 1)], lnA=np.log(1e8))
    Kc2 = np.exp(-(g[2] - g[0] - g[1])) * (101325.0 / (R * T)) ** -1
    cA, cB = 0.01, 0.03
    k2 = oracle.kinetics(m2, T, np.array([cA, cB, Kc2 * cA * cB]))
    assert abs(k2["r"][0]) < 1e-12 * k2["kf"][0] * cA * cB
    assert k2["uKeq"][0] == pytest.approx(1.0 / Kc2, rel=1e-12)


def test_third_body_efficiencies(oracle):
    """M = sum_k eff_k c_k (default efficiency 1), from the eff-1 storage."""
    m = _two_species_mech(reac=[(0, 1), (1, 1)], prod=[(2, 1)], lnA=np.log(2.0), reversible=False,
                          tb=[(0, 2.5), (2, 0.5)])
    c = np.array([0.2, 0.3, 0.4])
    M = 2.5 * 0.2 + 1.0 * 0.3 + 0.5 * 0.4
    assert oracle.kinetics(m, 1200.0, c)["r"][0] == pytest.approx(2.0 * M * 0.2 * 0.3, rel=1e-14)


def _falloff(troe, M_species_eff=None, single=-1):
    return _two_species_mech(reac=[(0, 1), (1, 1)], prod=[(2, 1)], lnA=np.log(1e12), beta=0.0, E_over_R=0.0,
                             reversible=False,
                             falloff=dict(inf=(np.log(1e8), 0.0, 0.0), troe=troe, single=single,
                                          eff=M_species_eff or []))


def test_lindemann_limits(oracle):
    m = _falloff(None)
    for scale in (1e-12, 1.0, 1e12):
        c = np.array([0.1, 0.2, 0.3]) * scale
        M = c.sum()
        k0, kinf = 1e12, 1e8
        Pr = k0 * M / kinf
        kmod = oracle.kinetics(m, 1000.0, c)["kmod"][0]
        assert kmod == pytest.approx(kinf * Pr / (1 + Pr), rel=1e-13)
    assert oracle.kinetics(m, 1000.0, np.array([1e-12, 1e-12, 1e-12]))["kmod"][0] == pytest.approx(
        1e12 * 3e-12, rel=1e-6)                                     # low-pressure limit k0 M
    assert oracle.kinetics(m, 1000.0, np.array([1e6, 1e6, 1e6]))["kmod"][0] == pytest.approx(1e8, rel=1e-6)


def test_troe_hand_value(oracle):
    """Fcent = 0.6 at Pr = 1, by hand: log10 Fc = -0.2218487, c = -0.4 - 0.67 log10 Fc = -0.2513613,
    n = 0.75 - 1.27 log10 Fc = 1.0317479, x = c / (n - 0.14 c) = -0.2513613 / 1.0669385 = -0.2355910,
    log10 F = -0.2218487 / (1 + x^2) = -0.2218487 / 1.0555031 = -0.2101830, F = 0.616335."""
    # Fcent = (1-a) e^{-T/T3} + a e^{-T/T1}: a = 0.6, T1 -> huge, T3 -> tiny
    m = _falloff((0.6, 1e-30, 1e30, 0.0))
    c = np.array([1e-4, 0.0, 0.0])
    c[1:] = 0.0
    M = c.sum()
    # choose concentrations so that Pr = k0 M / kinf = 1 -> M = 1e-4
    kmod = oracle.kinetics(m, 1000.0, c)["kmod"][0]
    F = kmod / (1e8 * 0.5)
    assert F == pytest.approx(0.616335, abs=2e-6)
    assert M == pytest.approx(1e-4)


def test_troe_limits(oracle):
    """Fcent = 1 reduces Troe to Lindemann; Pr -> 0 uses the asymptotic broadening."""
    m = _falloff((1.0, 1.0, 1e300, 0.0))                    # (1-a)=0, a e^{-T/1e300} = 1
    ml = _falloff(None)
    for c in (np.array([0.1, 0.2, 0.3]), np.array([1e-9, 0, 0])):
        assert oracle.kinetics(m, 1300.0, c)["kmod"][0] == pytest.approx(
            oracle.kinetics(ml, 1300.0, c)["kmod"][0], rel=1e-12)
    mt = _falloff((0.5, 200.0, 2000.0, 0.0))
    T = 1500.0
    Fc = 0.5 * np.exp(-T / 200.0) + 0.5 * np.exp(-T / 2000.0)
    kmod = oracle.kinetics(mt, T, np.zeros(3))["kmod"][0]    # Pr = 0 -> asymptote, k -> 0 * F
    assert kmod == 0.0
    c = np.array([1e-40, 0, 0])                              # Pr = 1e-36 < 1e-32
    kmod = oracle.kinetics(mt, T, c)["kmod"][0]
    F_asym = 10 ** (np.log10(Fc) / (1 + (1 / 0.14) ** 2))
    assert kmod / (1e12 * 1e-40) == pytest.approx(F_asym, rel=1e-9)


def test_single_collider(oracle):
    m = _falloff(None, single=1)
    c = np.array([0.1, 0.2, 0.3])
    M = 0.2 + 1e-16
    Pr = 1e12 * M / 1e8
    assert oracle.kinetics(m, 1000.0, c)["kmod"][0] == pytest.approx(1e8 * Pr / (1 + Pr), rel=1e-13)


@pytest.mark.parametrize("ns", [14, 53, 160])
def test_rates_match_dense_formulation(oracle, ns):
    m = workload_mechanism(ns)
    rng = np.random.default_rng(ns)
    for T in (650.0, 1400.0, 2300.0):
        x = rng.dirichlet(np.full(ns, 0.5))
        c = 101325.0 / (R * T) * x
        k = oracle.kinetics(m, T, c)
        r_ref, R_ref = dref.rates(m, T, c)
        scale = np.abs(r_ref).max()
        np.testing.assert_allclose(k["r"], r_ref, rtol=1e-9, atol=1e-12 * scale)
        np.testing.assert_allclose(k["R"], m.stoich_matrix() @ k["r"], rtol=1e-9, atol=1e-12 * scale)


@pytest.mark.parametrize("ns", [14, 53, 160])
def test_conservation_of_formation_rates(oracle, ns):
    """sum_k MW_k R_k = 0 (mass) and sum_k a_ek R_k = 0 (elements) to roundoff."""
    m = workload_mechanism(ns)
    rng = np.random.default_rng(7)
    for T in (500.0, 1500.0, 3000.0):
        c = 101325.0 / (R * T) * rng.dirichlet(np.full(ns, 0.5))
        Rf = oracle.kinetics(m, T, c)["R"]
        scale = np.abs(Rf).max() * m.mw.max()
        assert abs(m.mw @ Rf) <= 1e-12 * scale
        assert np.abs(m.elements.T @ Rf).max() <= 1e-12 * np.abs(Rf).max() * m.elements.max()


@pytest.mark.parametrize("ns", [14, 53])
def test_rhs_matches_dense_formulation(oracle, ns):
    m = workload_mechanism(ns)
    rng = np.random.default_rng(3)
    for T in (800.0, 2000.0):
        y = np.concatenate([rng.dirichlet(np.full(ns, 0.5)), [T]])
        f, qr = oracle.rhs(m, y)
        f_ref, qr_ref = dref.rhs(m, y)
        np.testing.assert_allclose(f, f_ref, rtol=1e-8, atol=1e-10 * np.abs(f_ref[:ns]).max())
        assert qr == pytest.approx(qr_ref, rel=1e-8)
        f0, _ = oracle.rhs(m, y, energy=False)
        assert f0[-1] == 0.0
        np.testing.assert_allclose(f0[:ns], f[:ns], rtol=1e-15)


def test_synthetic_mechanisms_are_element_balanced():
    for ns in (14, 53, 160):
        m = workload_mechanism(ns)
        assert np.abs(m.elements.T @ m.stoich_matrix()).max() == 0.0
        assert m.ns == ns
    m = synthetic_mechanism(14, 34, 4, 2, seed=1401)
    assert (m.lnA == workload_mechanism(14).lnA).all()      # seeded


def test_reverse_sums_is_a_rounding_variant(oracle):
    """The test option reverse_sums (formation rates summed over reactions in reverse order)
    changes only the rounding: on 14-species cells the step changes the final states (so the
    option is wired through) by less than the parity target of 10 x (rtol |y| + atol)."""
    from laminarsmoke_amd.fields import workload_cells
    mech, y = workload_cells(1, n=16)
    a, sa, _, ca = oracle.chemistry_step(mech, y, 1e-5, rtol=1e-7, atol=1e-12, threads=4)
    b, sb, _, cb = oracle.chemistry_step(mech, y, 1e-5, rtol=1e-7, atol=1e-12, threads=4, reverse_sums=True)
    assert np.array_equal(sa, sb)
    assert not np.array_equal(a, b)
    r = (np.abs(a - b) / (1e-7 * np.abs(a) + 1e-12)).max()
    assert r < 10.0, r
