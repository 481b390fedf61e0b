"""GPU: the reference's own H2/CO, H2 and CH4 skeletal (24 species, 155 reactions, one
SRI fall-off) mechanisms and its one-step global CH4 and H2 mechanisms (explicit FORD
orders) - CHEMKIN input read by laminarsmoke_amd.chemkin - through every kernel, against
the oracle.

* right-hand side and analytic Jacobian against the oracle (central differences of
  the oracle's right-hand side for the Jacobian);
* chemistry steps of premixed fuel/air cells (900-2200 K, half of them with a
  radical pool), dt = 1e-4 s so that many cells ignite within the step: team,
  plain and wide kernels (and the team kernel with 2 or 4 cells per wavefront) with the
  reference's forward-difference Jacobian under the
  envelope criterion with identical OdeStatus, and the analytic Jacobian (team and
  wide) against a tight-tolerance solution; element mass conserved.
"""
import os

import numpy as np
import pytest

from laminarsmoke_amd import api
from laminarsmoke_amd.chemkin import load_chemkin
from laminarsmoke_amd.fields import premixed_cells
from parity import (assert_parity, assert_parity_adaptive, assert_element_conservation, newton_matrix_changed,
                    oracle_envelope, oracle_jacobian_envelope, reconcile_statuses, ratio, split_method)
from test_gpu_jacobian import check_equations, check_jacobian

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RTOL, ATOL, DT = 1e-7, 1e-12, 1e-4


def mechanism(name):
    thermo = os.path.join(GOLDEN, name + "_thermo.CKT")
    if not os.path.exists(thermo):
        thermo = os.path.join(GOLDEN, "POLIMI_H2CO_1412_thermo.CKT")
    return load_chemkin(os.path.join(GOLDEN, name + ".CKI"), thermo)


def make_step(mech, method, jacobian):
    ode = api.ODE_Parameters()
    ode.SetRelativeTolerance(RTOL)
    ode.SetAbsoluteTolerance(ATOL)
    kw = split_method(method)
    kw["jacobian"] = jacobian
    return api.ChemistryStep(mech, ode, **kw)


@pytest.mark.parametrize("name", ["POLIMI_H2CO_1412", "POLIMI_H2_1412", "POLIMI_CH4_SKELETAL_1412", "GLOBAL_CH4_1step",
                                  "GLOBAL_H2_1step"])
def test_equations_and_jacobian(hip_device, oracle, name):
    mech = mechanism(name)
    y = premixed_cells(mech, 8, seed=3)
    print(f"{name}: equations, 8 cells", flush=True)
    step = make_step(mech, "team", "analytic")
    try:
        f, Ja = step.equations(y)
    finally:
        step.close()
    check_equations(oracle, mech, y, f, name)
    check_jacobian(oracle, mech, y, Ja, name)


@pytest.mark.parametrize("method", ["auto", "plain"])
def test_tutorial_settings(hip_device, oracle, method):
    """The MethaneCoflowFlame/02-skeletal-mechanism tutorial's OdeHomogeneous dictionary
    (rtol 1e-6, atol 1e-12, maximumOrder 5, fullPivoting true) on its mechanism: `auto`
    picks the team kernel (row pivoting), `plain` pivots fully like the oracle."""
    mech = mechanism("POLIMI_CH4_SKELETAL_1412")
    y = premixed_cells(mech, 48, seed=8)
    ode = api.ODE_Parameters()
    ode.SetupFromDictionary('OdeHomogeneous { odeSolver "OpenSMOKE"; relTolerance 1e-6; absTolerance 1e-12; '
                            'maximumOrder 5; fullPivoting true; }')
    print(f"tutorial settings, method={method}", flush=True)
    kw = dict(rtol=1e-6, atol=1e-12, full_pivoting=True)
    yo, st_o, *_ = oracle.chemistry_step(mech, y, DT, threads=8, **kw)
    env = oracle_envelope(oracle, mech, y, DT, 1e-6, 1e-12, yo, full_pivoting=True)
    s = api.ChemistryStep(mech, ode, method=method)
    try:
        used = s.launch()["method"]
        yg, st_g, *_ = s.integrate(y, DT)
    finally:
        s.close()
    assert used == ("team" if method == "auto" else "plain")
    assert np.array_equal(st_g, st_o)
    cell, _ = assert_parity(yg, yo, 1e-6, 1e-12, env, what=f"tutorial/{method}")
    print(f"tutorial/{used}: max ratio {cell.max():.3g} (envelope max {env.max():.3g})", flush=True)


@pytest.mark.parametrize("method,jac", [("team", "numerical"), ("plain", "numerical"), ("wide", "numerical"),
                                        ("team", "analytic"), ("wide", "analytic"), ("pack2", "numerical"),
                                        ("pack2", "analytic"), ("pack4", "numerical")])
@pytest.mark.parametrize("name", ["POLIMI_H2CO_1412", "POLIMI_H2_1412", "POLIMI_CH4_SKELETAL_1412", "GLOBAL_CH4_1step",
                                  "GLOBAL_H2_1step"])
def test_step(hip_device, oracle, name, method, jac):
    mech = mechanism(name)
    if method == "pack4" and mech.neq > 16:
        pytest.skip("4 cells per wavefront need ns+1 <= 16")
    y = premixed_cells(mech, 48, seed=5)
    print(f"{name}: step method={method} jacobian={jac} cells=48 dt={DT}", flush=True)
    yo, st_o, q_o, c_o = oracle.chemistry_step(mech, y, DT, rtol=RTOL, atol=ATOL, threads=8)
    step = make_step(mech, method, jac)
    try:
        yg, st_g, q_g, c_g = step.integrate(y, DT)
    finally:
        step.close()
    ignited = int((yo[mech.ns] - y[mech.ns] > 100.0).sum())
    variant = method + ("-analytic" if jac == "analytic" else "")
    env = oracle_envelope(oracle, mech, y, DT, RTOL, ATOL, yo)
    if newton_matrix_changed(variant):
        env = np.maximum(env, oracle_jacobian_envelope(oracle, mech, y, DT, RTOL, ATOL, yo))
        yo, st_o = reconcile_statuses(oracle, mech, y, DT, RTOL, ATOL, st_g, st_o, yo, what=f"{name}/{variant}")
    assert np.array_equal(st_g, st_o), (st_g, st_o)
    (cell, frac), env = assert_parity_adaptive(oracle, mech, y, DT, RTOL, ATOL, yo, yg, env, what=f"{name}/{variant}")
    assert_element_conservation(mech, y, yg, c_g[:, 10], what=f"{name}/{method}/{jac}", y_ref=yo)
    np.testing.assert_allclose(yg[:mech.ns].sum(0), 1.0, atol=1e-14)
    print(f"{name}/{method}/{jac}: max ratio {cell.max():.3g}; {ignited} of 48 cells ignite; steps gpu/oracle "
          f"median {np.median(c_g[:, 0])}/{np.median(c_o[:, 0])}, max dT {np.abs(yg[-1] - yo[-1]).max():.3g} K",
          flush=True)
