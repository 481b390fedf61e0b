"""Oracle: the constant-volume reactor and the Strang "compact" output of the cell loop.

* Energy invariants of the right-hand sides, from thermodynamics alone: a constant-
  pressure adiabatic reactor keeps its mixture enthalpy, sum_k h_k dY_k/dt + cp dT/dt = 0;
  a constant-volume one its internal energy, sum_k u_k dY_k/dt + cv dT/dt = 0
  (u_k = h_k - R T / MW_k, cv = cp - R / MW).  A dropped or mis-signed term of either
  temperature equation fails these.
* At the ideal-gas density of the state the two reactors have the same species rows.
* A constant-volume chemistry step conserves element mass and internal energy, and a
  cell that ignites raises its pressure rho R T / MW.
* Compact output: rho_f (Y_f - Y_in) / dt and rho_f cp_f (T_f - T_in) / dt of the
  ordinary step's final state (rho_f, cp_f at the reactor pressure), zero for
  dt <= 1e-14, RT zero without the energy equation.
"""
import os

import numpy as np
import pytest

from laminarsmoke_amd.chemkin import load_chemkin
from laminarsmoke_amd.fields import premixed_cells, workload_cells

R = 8314.4621
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def h2co():
    return load_chemkin(os.path.join(GOLDEN, "POLIMI_H2CO_1412.CKI"), os.path.join(GOLDEN, "POLIMI_H2CO_1412_thermo.CKT"))


def cases():
    m1, y1 = workload_cells(1, n=6)
    m2 = h2co()
    return [(m1, y1), (m2, premixed_cells(m2, 6, seed=3))]


def mix(oracle, mech, y):
    """Per-mass h_k, u_k [J/kg], cp, cv [J/kg/K], MW of one state."""
    T = y[-1]
    cp, h, _ = oracle.species_thermo(mech, T)
    hk = h * R * T / mech.mw
    uk = (h - 1.0) * R * T / mech.mw
    Y = y[:-1]
    return hk, uk, np.sum(Y * cp / mech.mw) * R, np.sum(Y * (cp - 1.0) / mech.mw) * R, 1.0 / np.sum(Y / mech.mw)


@pytest.mark.parametrize("k", [0, 1])
def test_energy_invariants(oracle, k):
    mech, y = cases()[k]
    for c in range(y.shape[1]):
        yc = y[:, c]
        hk, uk, cp, cv, MW = mix(oracle, mech, yc)
        rho = 101325.0 * MW / (R * yc[-1])
        fp, _ = oracle.rhs(mech, yc)
        fv, _ = oracle.rhs(mech, yc, rho0=rho)
        scale = np.abs(hk * fp[:-1]).sum() + 1e-300
        assert abs(np.sum(hk * fp[:-1]) + cp * fp[-1]) <= 1e-12 * scale
        assert abs(np.sum(uk * fv[:-1]) + cv * fv[-1]) <= 1e-12 * (np.abs(uk * fv[:-1]).sum() + 1e-300)
        np.testing.assert_allclose(fv[:-1], fp[:-1], rtol=1e-13, atol=1e-13 * np.abs(fp[:-1]).max())


def test_constant_volume_step_conserves(oracle):
    mech = h2co()
    y = premixed_cells(mech, 16, seed=9, T_lo=1100.0, T_hi=1600.0)
    yo, st, q, cnt = oracle.chemistry_step(mech, y, 1e-3, rtol=1e-9, atol=1e-15, constant_volume=True, threads=8)
    assert (st == 2).all()
    em = mech.element_mass_matrix()
    np.testing.assert_allclose(em @ yo[:-1], em @ y[:-1], atol=1e-12)
    rose = 0
    for c in range(y.shape[1]):
        _, uk0, _, cv0, MW0 = mix(oracle, mech, y[:, c])
        _, ukf, _, cvf, MWf = mix(oracle, mech, yo[:, c])
        u0, uf = np.sum(y[:-1, c] * uk0), np.sum(yo[:-1, c] * ukf)
        assert abs(uf - u0) <= 1e-6 * cv0 * y[-1, c]
        if yo[-1, c] > y[-1, c] + 500.0:
            rose += 1
            P0 = 101325.0
            Pf = (P0 * MW0 / (R * y[-1, c])) * R * yo[-1, c] / MWf
            assert Pf > 1.3 * P0
    assert rose >= 4


def test_density_argument(oracle):
    mech, y = cases()[1]
    MW = 1.0 / np.sum(y[:-1] / mech.mw[:, None], axis=0)
    rho = 101325.0 * MW / (R * y[-1])
    a, sa, *_ = oracle.chemistry_step(mech, y, 1e-4, constant_volume=True)
    b, sb, *_ = oracle.chemistry_step(mech, y, 1e-4, constant_volume=True, rho=rho)
    c, sc, *_ = oracle.chemistry_step(mech, y, 1e-4, constant_volume=True, rho=1.5 * rho)
    assert (sa == 2).all() and (sb == 2).all() and (sc == 2).all()
    np.testing.assert_allclose(b, a, rtol=1e-6, atol=1e-10)
    assert np.abs(c[-1] - a[-1]).max() > 1.0          # a denser mixture reacts differently


@pytest.mark.parametrize("cv", [False, True])
@pytest.mark.parametrize("energy", [True, False])
def test_compact_sources(oracle, cv, energy):
    mech, y = cases()[1]
    dt = np.full(y.shape[1], 1e-4)
    dt[-1] = 1e-15
    yf, st, *_ = oracle.chemistry_step(mech, y, dt, constant_volume=cv, energy=energy)
    rr, st2, *_ = oracle.chemistry_step(mech, y, dt, constant_volume=cv, energy=energy, compact=True)
    assert np.array_equal(st, st2)
    for c in range(y.shape[1]):
        _, _, cp, _, MW = mix(oracle, mech, yf[:, c])
        rho = 101325.0 * MW / (R * yf[-1, c])
        if dt[c] <= 1e-14:
            assert (rr[:, c] == 0.0).all()
            continue
        want = rho * (yf[:-1, c] - y[:-1, c]) / dt[c]
        np.testing.assert_allclose(rr[:-1, c], want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
        rt = rho * cp * (yf[-1, c] - y[-1, c]) / dt[c] if energy else 0.0
        assert rr[-1, c] == pytest.approx(rt, rel=1e-12, abs=1e-300)
