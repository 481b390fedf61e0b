"""GPU parity: the HIP chemistry step against the CPU oracle on the same seeded cells.

Criteria (tests/parity.py): identical OdeStatus per cell; per-cell deviation within
max(10, 3 x oracle self-envelope) x (rtol*|y| + atol); element mass conserved to 1e-12
(scaled by clamp activity); sum Y = 1.
"""
import numpy as np
import pytest

from laminarsmoke_amd import api
from laminarsmoke_amd.fields import workload_cells, coflow_field
from laminarsmoke_amd.mechanism import workload_mechanism
from parity import assert_parity, assert_element_conservation, oracle_envelope

pytestmark = pytest.mark.gpu

RTOL, ATOL, DT = 1e-7, 1e-12, 1e-5


def gpu_step(mech, y, dt, method, rtol=RTOL, atol=ATOL, **kw):
    ode = api.ODE_Parameters()
    ode.SetRelativeTolerance(rtol)
    ode.SetAbsoluteTolerance(atol)
    jacobian = "numerical"
    if method.endswith("-analytic"):
        method, jacobian = method[:-len("-analytic")], "analytic"
    step = api.ChemistryStep(mech, ode, method=method, jacobian=jacobian, **kw)
    try:
        return step.integrate(y, dt)
    finally:
        step.close()


def _check(oracle, mech, y, method, what):
    yo, st_o, q_o, c_o = oracle.chemistry_step(mech, y, DT, rtol=RTOL, atol=ATOL, threads=8)
    env = oracle_envelope(oracle, mech, y, DT, RTOL, ATOL, yo)
    yg, st_g, q_g, c_g = gpu_step(mech, y, DT, method)
    assert np.array_equal(st_g, st_o), (st_g, st_o)
    cell, frac = assert_parity(yg, yo, RTOL, ATOL, env, what=what)
    assert_element_conservation(mech, y, yg, c_g[:, 10], what=what)
    np.testing.assert_allclose(yg[:mech.ns].sum(0), 1.0, atol=1e-14)
    print(f"{what}: max ratio {cell.max():.3g} (envelope max {env.max():.3g}), "
          f"components within 10x: {100 * frac:.2f}%, steps gpu/oracle median {np.median(c_g[:, 0])}/"
          f"{np.median(c_o[:, 0])}", flush=True)
    return yg, yo, q_g, q_o


# The wide kernel with the analytic Jacobian (FP32 dR/dc and FP32 Newton LU) is held
# to the tight-solution criterion instead (tests/test_gpu_jacobian.py).
@pytest.mark.parametrize("method", ["team", "plain", "wide", "team-analytic"])
def test_cfg1_parity(hip_device, oracle, method):
    mech, y = workload_cells(1, n=64)
    print(f"cfg1 parity method={method} cells={y.shape[1]}", flush=True)
    yg, yo, q_g, q_o = _check(oracle, mech, y, method, f"cfg1/{method}")
    # heat release of the final state, same cells
    np.testing.assert_allclose(q_g, q_o, rtol=1e-3)


@pytest.mark.parametrize("method", ["team", "plain", "wide", "team-analytic"])
def test_cfg2_sample_parity(hip_device, oracle, method):
    mech = workload_mechanism(53)
    yfull, flame = coflow_field(mech, 512, 512, 22)
    idx = np.concatenate([np.where(flame)[0][:6], np.where(~flame)[0][:10]])
    y = np.ascontiguousarray(yfull[:, idx])
    print(f"cfg2 parity method={method} cells={y.shape[1]}", flush=True)
    _check(oracle, mech, y, method, f"cfg2/{method}")
