"""GPU: reactor equations and the analytic Jacobian.

* lsm_equations' right-hand side against the oracle's (GetEquations of the
  constant-pressure batch reactor) on flame and cold cells;
* the analytic Jacobian against central differences of the oracle's right-hand
  side (the oracle is the independent function here; nothing of the analytic
  formulas is restated in the test);
* a hand-built mechanism whose reactions take the general stoichiometry path of
  the team kernel (fractional orders, four reactants, a single-collider Lindemann
  fall-off, negative A, irreversible steps): equations, Jacobian and a full
  chemistry step against the oracle;
* chemistry steps with the analytic Jacobian against a tight-tolerance solution:
  no larger error than the reference's forward-difference Jacobian gives.
"""
import numpy as np
import pytest

from laminarsmoke_amd import api
from laminarsmoke_amd.fields import workload_cells, coflow_field
from laminarsmoke_amd.mechanism import workload_mechanism, custom_mechanism
from parity import assert_parity, oracle_envelope, ratio, check_against_truth

pytestmark = pytest.mark.gpu

RTOL, ATOL, DT = 1e-7, 1e-12, 1e-5


def make_step(mech, method, jacobian="numerical"):
    ode = api.ODE_Parameters()
    ode.SetRelativeTolerance(RTOL)
    ode.SetAbsoluteTolerance(ATOL)
    return api.ChemistryStep(mech, ode, method=method, jacobian=jacobian)


def general_mechanism():
    """Seven species (thermochemistry of the 14-species mechanism's first seven) with
    reactions outside the fixed three-by-three format."""
    base = workload_mechanism(14)
    sp = base.species[:7]
    nasa = base.nasa[:7]
    R = [
        dict(reac=[(0, 0.5), (1, 1.5)], prod=[(2, 1.0)], lnA=np.log(3e7), beta=0.3, E_over_R=4000.0,
             reversible=False),
        dict(reac=[(0, 1), (1, 1), (3, 1), (4, 1)], prod=[(5, 1), (6, 1)], lnA=np.log(2e9), beta=-0.5,
             E_over_R=2000.0),
        dict(reac=[(2, 2)], prod=[(3, 1), (4, 1)], lnA=np.log(5e8), beta=0.0, E_over_R=9000.0,
             tb=[(0, 2.5), (5, 6.0)]),
        dict(reac=[(3, 1), (4, 1)], prod=[(5, 1)], lnA=np.log(1e9), beta=-1.0, E_over_R=500.0,
             falloff=dict(inf=(np.log(4e7), 0.2, 1500.0), troe=None, single=6)),
        dict(reac=[(5, 1)], prod=[(3, 1), (6, 1)], lnA=np.log(2e3), beta=0.5, E_over_R=7000.0,
             negative_A=True, reversible=False),
        dict(reac=[(6, 2), (0, 1)], prod=[(1, 2), (2, 1)], lnA=np.log(7e10), beta=0.1, E_over_R=6000.0),
        dict(reac=[(1, 1), (2, 1)], prod=[(3, 1)], lnA=np.log(3e6), beta=-0.7, E_over_R=800.0,
             falloff=dict(inf=(np.log(6e4), 0.1, 300.0), troe=(0.6, 200.0, 1500.0, 4000.0),
                          eff=[(0, 3.0), (4, 0.5)])),
    ]
    # Shift each species' entropy constant (a7 of both ranges) by the least-norm
    # amount that puts every reversible reaction at K = 1 at 1300 K, so that the
    # reverse rates stay commensurate with the forward ones.
    Ts = 1300.0
    a = nasa[:, :7]
    g = (a[:, 0] * (1 - np.log(Ts)) - a[:, 1] * Ts / 2 - a[:, 2] * Ts ** 2 / 6 - a[:, 3] * Ts ** 3 / 12
         - a[:, 4] * Ts ** 4 / 20 + a[:, 5] / Ts - a[:, 6])
    rows, rhs = [], []
    for r in R:
        if not r.get("reversible", True):
            continue
        nu = np.zeros(len(sp))
        for k, v in r["reac"]:
            nu[k] -= v
        for k, v in r["prod"]:
            nu[k] += v
        rows.append(nu)
        rhs.append(nu @ g - nu.sum() * np.log(101325.0 / 8314.4621 / Ts))
    delta = np.linalg.lstsq(np.array(rows), np.array(rhs), rcond=None)[0]
    nasa = nasa.copy()
    nasa[:, 6] += delta
    nasa[:, 13] += delta
    return custom_mechanism(sp, nasa, R, name="general-7sp")


def general_states(mech, n, seed=7):
    rng = np.random.default_rng(seed)
    y = np.empty((mech.neq, n))
    y[:mech.ns] = rng.uniform(0.05, 1.0, (mech.ns, n))
    y[:mech.ns] /= y[:mech.ns].sum(0)
    y[mech.ns] = rng.uniform(900.0, 1800.0, n)
    return y


def cd_jacobian(oracle, mech, y):
    """Central differences of the oracle's right-hand side at one state y[neq]."""
    neq = y.size
    J = np.empty((neq, neq))
    for j in range(neq):
        h = 1e-5 * y[j] if j == mech.ns else 1e-6 * max(abs(y[j]), 1e-3)
        yp, ym = y.copy(), y.copy()
        yp[j] += h
        ym[j] -= h
        fp, _ = oracle.rhs(mech, yp)
        fm, _ = oracle.rhs(mech, ym)
        J[:, j] = (fp - fm) / (2.0 * h)
    return J


def check_jacobian(oracle, mech, y, Ja, what):
    worst = 0.0
    for c in range(y.shape[1]):
        Jc = cd_jacobian(oracle, mech, y[:, c])
        scale = np.abs(Jc).max(axis=1, keepdims=True) + 1e-300
        err = np.abs(Ja[c] - Jc) / scale
        worst = max(worst, float(err.max()))
        assert err.max() < 1e-5, (f"{what}: cell {c} worst entry {np.unravel_index(err.argmax(), err.shape)} "
                                  f"rel err {err.max():.3g}")
    print(f"{what}: analytic vs central-difference Jacobian, worst row-scaled error {worst:.3g}", flush=True)


def check_equations(oracle, mech, y, f, what):
    """Species rows scaled by the largest species rate of the cell, the temperature
    row by itself (formation rates are sums of terms of either sign, so a rate far
    below the cell's largest carries that sum's rounding)."""
    fo = np.stack([oracle.rhs(mech, y[:, c])[0] for c in range(y.shape[1])], axis=1)
    scale = np.empty_like(fo)
    scale[:mech.ns] = np.abs(fo[:mech.ns]).max(axis=0, keepdims=True)
    scale[mech.ns] = np.abs(fo[mech.ns])
    err = np.abs(f - fo) / (scale + 1e-300)
    assert err.max() < 1e-11, f"{what}: rhs rel err {err.max():.3g}"
    print(f"{what}: rhs vs oracle, worst error {err.max():.3g}", flush=True)


def states(cfg):
    if cfg == 1:
        mech, y = workload_cells(1, n=8)
        return mech, y
    mech = workload_mechanism(53)
    yfull, flame = coflow_field(mech, 512, 512, 22)
    idx = np.concatenate([np.where(flame)[0][:5], np.where(~flame)[0][:2]])
    return mech, np.ascontiguousarray(yfull[:, idx])


@pytest.mark.parametrize("cfg", [1, 2])
def test_equations_and_jacobian(hip_device, oracle, cfg):
    mech, y = states(cfg)
    print(f"equations cfg{cfg} cells={y.shape[1]}", flush=True)
    step = make_step(mech, "team")
    try:
        f, Ja = step.equations(y)
    finally:
        step.close()
    check_equations(oracle, mech, y, f, f"cfg{cfg}")
    check_jacobian(oracle, mech, y, Ja, f"cfg{cfg}")


def test_general_mechanism_equations_and_jacobian(hip_device, oracle):
    mech = general_mechanism()
    y = general_states(mech, 6)
    print("equations general-7sp cells=6", flush=True)
    step = make_step(mech, "team")
    try:
        f, Ja = step.equations(y)
    finally:
        step.close()
    check_equations(oracle, mech, y, f, "general")
    check_jacobian(oracle, mech, y, Ja, "general")


@pytest.mark.parametrize("method,jac", [("team", "numerical"), ("plain", "numerical"), ("team", "analytic"),
                                        ("wide", "numerical"), ("wide", "analytic")])
def test_general_mechanism_step(hip_device, oracle, method, jac):
    mech = general_mechanism()
    y = general_states(mech, 32, seed=11)
    print(f"general-7sp step method={method} jacobian={jac}", flush=True)
    yo, st_o, _, _ = oracle.chemistry_step(mech, y, DT, rtol=RTOL, atol=ATOL, threads=8)
    step = make_step(mech, method, jac)
    try:
        yg, st_g, _, _ = step.integrate(y, DT)
    finally:
        step.close()
    if jac == "numerical":
        assert np.array_equal(st_g, st_o), (st_g, st_o)
        env = oracle_envelope(oracle, mech, y, DT, RTOL, ATOL, yo)
        cell, _ = assert_parity(yg, yo, RTOL, ATOL, env, what=f"general/{method}")
    else:
        truth, *_ = oracle.chemistry_step(mech, y, DT, rtol=1e-12, atol=1e-20, threads=8)
        check_against_truth(yo, st_o, yg, st_g, truth, RTOL, ATOL, f"general/{method}/analytic")
        cell = ratio(yg, yo, RTOL, ATOL).max(axis=0)
    print(f"general/{method}/{jac}: max ratio {cell.max():.3g}", flush=True)


@pytest.mark.parametrize("method", ["team", "wide"])
@pytest.mark.parametrize("cfg", [1, 2])
def test_analytic_jacobian_step_accuracy(hip_device, oracle, cfg, method):
    if cfg == 1:
        mech, y = workload_cells(1, n=64)
    else:
        mech = workload_mechanism(53)
        yfull, flame = coflow_field(mech, 512, 512, 22)
        idx = np.concatenate([np.where(flame)[0][:12], np.where(~flame)[0][:4]])
        y = np.ascontiguousarray(yfull[:, idx])
    print(f"analytic-Jacobian step cfg{cfg} method={method} cells={y.shape[1]}", flush=True)
    yo, st_o, _, c_o = oracle.chemistry_step(mech, y, DT, rtol=RTOL, atol=ATOL, threads=8)
    truth, st_t, *_ = oracle.chemistry_step(mech, y, DT, rtol=1e-12, atol=1e-20, threads=8)
    step = make_step(mech, method, "analytic")
    try:
        yg, st_g, _, c_g = step.integrate(y, DT)
    finally:
        step.close()
    check_against_truth(yo, st_o, yg, st_g, truth, RTOL, ATOL, f"cfg{cfg}/{method}/analytic")
    print(f"cfg{cfg}/{method}/analytic: steps {np.median(c_g[:, 0])} vs {np.median(c_o[:, 0])}, rhs per cell "
          f"{c_g[:, 9].mean():.1f} vs {c_o[:, 9].mean():.1f}", flush=True)
    assert (c_g[:, 2] == 0).all()          # no RHS evaluations spent on Jacobians
    np.testing.assert_allclose(yg[:mech.ns].sum(0), 1.0, atol=1e-14)
