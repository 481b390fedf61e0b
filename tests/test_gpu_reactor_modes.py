"""GPU: the constant-volume reactor and the Strang "compact" output.

* constant-volume chemistry steps (team with 1 or 4 cells per wavefront, wide, plain; the reference's
  forward-difference Jacobian) against the oracle: identical OdeStatus and the
  envelope criterion, on the 14-species workload cells and on the reference's H2/CO
  mechanism (dt = 1e-4 s, most cells ignite);
* per-cell densities passed by the caller (rhoCells) against the oracle given the same
  densities;
* compact output: the sources each kernel writes equal rho_f (Y_f - Y_in)/dt and
  rho_f cp_f (T_f - T_in)/dt of its own ordinary step, with the same launch (the
  ordinary step's parity is the test above and test_gpu_parity);
* the analytic Jacobian is refused for a constant-volume handle.
"""
import os

import numpy as np
import pytest

from laminarsmoke_amd import api
from laminarsmoke_amd.chemkin import load_chemkin
from laminarsmoke_amd.fields import premixed_cells, workload_cells
from parity import assert_parity, assert_element_conservation, oracle_envelope, split_method

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RTOL, ATOL = 1e-7, 1e-12
R = 8314.4621


def cells(which):
    if which == "cfg1":
        mech, y = workload_cells(1, n=64)
        return mech, y, 1e-5
    mech = load_chemkin(os.path.join(GOLDEN, "POLIMI_H2CO_1412.CKI"),
                        os.path.join(GOLDEN, "POLIMI_H2CO_1412_thermo.CKT"))
    return mech, premixed_cells(mech, 48, seed=5), 1e-4


def step(mech, method, **kw):
    ode = api.ODE_Parameters()
    ode.SetRelativeTolerance(RTOL)
    ode.SetAbsoluteTolerance(ATOL)
    return api.ChemistryStep(mech, ode, **{**split_method(method), **kw})


@pytest.mark.parametrize("method", ["team", "wide", "plain", "pack4"])
@pytest.mark.parametrize("which", ["cfg1", "h2co"])
def test_constant_volume_parity(hip_device, oracle, which, method):
    mech, y, dt = cells(which)
    print(f"constant volume {which} method={method} cells={y.shape[1]} dt={dt}", flush=True)
    yo, st_o, q_o, c_o = oracle.chemistry_step(mech, y, dt, rtol=RTOL, atol=ATOL, constant_volume=True, threads=8)
    env = oracle_envelope(oracle, mech, y, dt, RTOL, ATOL, yo, constant_volume=True)
    s = step(mech, method, constant_volume=True)
    try:
        yg, st_g, q_g, c_g = s.integrate(y, dt)
    finally:
        s.close()
    assert np.array_equal(st_g, st_o), (st_g, st_o)
    cell, _ = assert_parity(yg, yo, RTOL, ATOL, env, what=f"cv/{which}/{method}")
    assert_element_conservation(mech, y, yg, c_g[:, 10], what=f"cv/{which}/{method}", y_ref=yo)
    np.testing.assert_allclose(q_g, q_o, rtol=1e-3, atol=1e-6 * np.abs(q_o).max())
    print(f"cv/{which}/{method}: max ratio {cell.max():.3g} (envelope max {env.max():.3g}), T rise up to "
          f"{(yg[-1] - y[-1]).max():.0f} K", flush=True)


def test_cell_densities(hip_device, oracle):
    mech, y, dt = cells("h2co")
    MW = 1.0 / np.sum(y[:-1] / mech.mw[:, None], axis=0)
    rho = 1.3 * 101325.0 * MW / (R * y[-1])
    print("constant volume with caller densities (1.3 x ideal gas)", flush=True)
    yo, st_o, *_ = oracle.chemistry_step(mech, y, dt, rtol=RTOL, atol=ATOL, constant_volume=True, rho=rho, threads=8)
    env = oracle_envelope(oracle, mech, y, dt, RTOL, ATOL, yo, constant_volume=True, rho=rho)
    for method in ("team", "wide", "plain"):
        s = step(mech, method, constant_volume=True)
        try:
            yg, st_g, *_ = s.integrate(y, dt, rho=rho)
        finally:
            s.close()
        assert np.array_equal(st_g, st_o)
        cell, _ = assert_parity(yg, yo, RTOL, ATOL, env, what=f"rho/{method}")
        print(f"rho/{method}: max ratio {cell.max():.3g}", flush=True)


@pytest.mark.parametrize("cv", [False, True])
@pytest.mark.parametrize("method", ["team", "wide", "plain", "pack4"])
def test_compact_output(hip_device, oracle, method, cv):
    mech, y, _ = cells("h2co")
    n = y.shape[1]
    dt = np.full(n, 1e-4)
    dt[-1] = 1e-15
    print(f"compact output method={method} constant_volume={cv}", flush=True)
    s = step(mech, method, constant_volume=cv)
    sc = step(mech, method, constant_volume=cv, compact=True)
    try:
        yf, st, *_ = s.integrate(y, dt)
        rr, st2, *_ = sc.integrate(y, dt)
    finally:
        s.close()
        sc.close()
    assert np.array_equal(st, st2)
    MW = 1.0 / np.sum(yf[:-1] / mech.mw[:, None], axis=0)
    rho = 101325.0 * MW / (R * yf[-1])
    cp = np.array([np.sum(yf[:-1, c] * oracle.species_thermo(mech, yf[-1, c])[0] / mech.mw) * R for c in range(n)])
    want = rho * (yf[:-1] - y[:-1]) / dt
    want_t = rho * cp * (yf[-1] - y[-1]) / dt
    want[:, -1] = 0.0
    want_t[-1] = 0.0
    np.testing.assert_allclose(rr[:-1], want, rtol=1e-11, atol=1e-11 * np.abs(want).max())
    np.testing.assert_allclose(rr[-1], want_t, rtol=1e-11, atol=1e-11 * np.abs(want_t).max())
    print(f"compact/{method}/cv={cv}: RR up to {np.abs(want).max():.3g} kg/m3/s, RT up to "
          f"{np.abs(want_t).max():.3g} W/m3", flush=True)


def test_analytic_jacobian_refused(hip_device):
    mech, y, dt = cells("cfg1")
    with pytest.raises(api.LsmError):
        step(mech, "team", constant_volume=True, jacobian="analytic")
