"""GPU: explicitly reversible (REV), chemically activated (CABR), pressure-logarithmic
(PLOG) and Chebyshev reactions through every kernel, against the oracle, at pressures on, between
and outside the PLOG levels, and in the constant-volume reactor (whose kinetics see
the pressure cTot R T of the current state).

The mechanism is the reference's H2/CO mechanism (POLIMI_H2CO_1412) with
* explicit reverse Arrhenius parameters on six reversible reactions (one of them a
  third-body reaction, one also PLOG): least-squares fits of k/Kc over 700-2600 K, so
  that the chemistry stays close to its thermodynamics;
* its H+O2(+M) Troe and CO+O(+M) Lindemann fall-off reactions turned into chemically
  activated ones (Troe, and SRI with d and e), with k0 = the fall-off k_inf and
  k_inf = k_inf^2/k0 so that Pr is the fall-off's;
* PLOG tables of four levels (0.1-100 atm) on four reactions (reversible,
  irreversible, and with REV);
* Chebyshev tables (6 x 4 over 250-4000 K and 0.005-1000 atm) on two reactions
  (reversible and irreversible): least-squares fits of their Arrhenius law times a
  pressure factor.
"""
import dataclasses
import os

import numpy as np
import pytest

from laminarsmoke_amd import api
from laminarsmoke_amd.chemkin import load_chemkin
from laminarsmoke_amd.fields import premixed_cells
from laminarsmoke_amd.mechanism import FALLOFF_CHEMACT, FALLOFF_SRI, FALLOFF_TROE, RXN_EXPLICIT_REV
from parity import (assert_element_conservation, assert_parity, assert_parity_adaptive, newton_matrix_changed,
                    oracle_envelope, oracle_jacobian_envelope, reconcile_statuses, ratio)

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RTOL, ATOL, DT = 1e-7, 1e-12, 1e-4
ATM = 101325.0


def _find(m, reac, prod):
    key = (sorted(reac), sorted(prod))
    hits = []
    for j in range(m.nr):
        r = sorted(m.species[s] for s in m.reac_sp[m.reac_ptr[j]:m.reac_ptr[j + 1]])
        p = sorted(m.species[s] for s in m.prod_sp[m.prod_ptr[j]:m.prod_ptr[j + 1]])
        if (r, p) == key:
            hits.append(j)
    assert hits, key
    return hits[0]


def forms_mechanism(oracle):
    base = load_chemkin(os.path.join(GOLDEN, "POLIMI_H2CO_1412.CKI"),
                        os.path.join(GOLDEN, "POLIMI_H2CO_1412_thermo.CKT"))
    rflags, rev_arr = base.rflags.copy(), base.rev_arr.copy()
    Ts = np.linspace(700.0, 2600.0, 40)
    c = np.full(base.ns, 1e-3)
    ks = [oracle.kinetics(base, T, c) for T in Ts]
    X = np.stack([np.ones_like(Ts), np.log(Ts), -1.0 / Ts], axis=1)
    rev = [_find(base, ["H", "O2"], ["OH", "O"]), _find(base, ["O", "H2"], ["OH", "H"]),
           _find(base, ["OH"], ["O", "H2O"]), _find(base, ["H", "OH"], ["H2O"]),
           _find(base, ["CO", "OH"], ["CO2", "H"]), _find(base, ["H", "H2O"], ["H2", "OH"])]
    for j in rev:
        lnkr = np.log([k["kf"][j] * k["uKeq"][j] for k in ks])
        a, b, e = np.linalg.lstsq(X, lnkr, rcond=None)[0]
        rflags[j] |= RXN_EXPLICIT_REV
        rev_arr[j] = (a, b, e)
    fo_type, fo_inf, fo_sri = base.fo_type.copy(), base.fo_inf.copy(), base.fo_sri.copy()
    lnA, beta, EoR = base.lnA.copy(), base.beta.copy(), base.E_over_R.copy()
    for jr, form, sri in ((_find(base, ["H", "O2"], ["HO2"]), FALLOFF_TROE, None),
                          (_find(base, ["CO", "O"], ["CO2"]), FALLOFF_SRI, (0.5, 200.0, 900.0, 1.2, 0.1))):
        s = int(np.where(base.fo_rxn == jr)[0][0])
        k0 = np.array([lnA[jr], beta[jr], EoR[jr]])
        kinf = fo_inf[s].copy()
        lnA[jr], beta[jr], EoR[jr] = kinf
        fo_inf[s] = 2.0 * kinf - k0
        fo_type[s] = form | FALLOFF_CHEMACT
        if sri is not None:
            fo_sri[s] = sri
    pl = [_find(base, ["H", "HO2"], ["OH"]), _find(base, ["HO2"], ["H2O2", "O2"]),
          _find(base, ["HCO", "HO2"], ["H", "OH", "CO2"]), rev[1]]
    levels = []
    for j in pl:
        lev = []
        for P, f, db, fe in ((0.1, 0.5, 0.05, 0.95), (1.0, 1.0, 0.0, 1.0), (10.0, 1.8, -0.05, 1.05),
                             (100.0, 2.5, -0.1, 1.1)):
            lev.append((P * ATM, lnA[j] + np.log(f), beta[j] + db, EoR[j] * fe))
        levels.append(lev)
    ch = [_find(base, ["H", "HO2"], ["H2", "O2"]), _find(base, ["HO2", "H2O"], ["H2O2", "OH"])]
    Tl, Pl = (250.0, 4000.0), (0.005 * ATM, 1000.0 * ATM)
    xs = np.cos(np.pi * (np.arange(24) + 0.5) / 24)                 # Chebyshev nodes in [-1, 1]
    Tn = 2.0 / ((xs + 1.0) * (1.0 / Tl[1] - 1.0 / Tl[0]) + 2.0 / Tl[0])
    Pn = 10 ** ((xs[:12] * (np.log10(Pl[1]) - np.log10(Pl[0])) + np.log10(Pl[1]) + np.log10(Pl[0])) / 2.0)
    xx, yy = np.meshgrid(xs, xs[:12], indexing="ij")
    TT, PP = np.meshgrid(Tn, Pn, indexing="ij")
    V = np.polynomial.chebyshev.chebvander2d(xx.ravel(), yy.ravel(), [5, 3])
    ch_a = []
    for j in ch:
        target = (lnA[j] + beta[j] * np.log(TT) - EoR[j] / TT) / np.log(10.0) + \
            0.15 * np.log10(PP / ATM) * (1.0 - 300.0 / TT)
        ch_a.append(np.linalg.lstsq(V, target.ravel(), rcond=None)[0])
    m = dataclasses.replace(base, rflags=rflags, rev_arr=rev_arr, lnA=lnA, beta=beta, E_over_R=EoR,
                            fo_type=fo_type, fo_inf=fo_inf, fo_sri=fo_sri, pl_rxn=np.array(pl, dtype=np.int32),
                            pl_ptr=np.array([0, 4, 8, 12, 16], dtype=np.int32),
                            pl_par=np.array([v for x in levels for v in x]),
                            ch_rxn=np.array(ch, dtype=np.int32), ch_dims=np.array([[6, 4], [6, 4]], dtype=np.int32),
                            ch_lim=np.array([[*Tl, *Pl, 1.0]] * 2), ch_ptr=np.array([0, 24, 48], dtype=np.int32),
                            ch_a=np.concatenate(ch_a), name="H2CO_rev_cabr_plog_cheb", _struct=None)
    return m


@pytest.fixture(scope="module")
def mech(oracle):
    return forms_mechanism(oracle)


def _rhs(oracle, mech, y, P):
    return np.stack([oracle.rhs(mech, y[:, c], P=P)[0] for c in range(y.shape[1])], axis=1)


def _cd_jacobian(oracle, mech, y, P):
    neq = y.size
    J = np.empty((neq, neq))
    for j in range(neq):
        h = 1e-5 * y[j] if j == mech.ns else 1e-6 * max(abs(y[j]), 1e-3)
        yp, ym = y.copy(), y.copy()
        yp[j] += h
        ym[j] -= h
        J[:, j] = (oracle.rhs(mech, yp, P=P)[0] - oracle.rhs(mech, ym, P=P)[0]) / (2.0 * h)
    return J


PRESSURES = [1.0 * ATM, 3.7 * ATM, 0.02 * ATM, 300.0 * ATM]


@pytest.mark.parametrize("P", PRESSURES)
def test_equations_and_jacobian(hip_device, oracle, mech, P):
    y = premixed_cells(mech, 6, seed=11)
    print(f"equations at P = {P / ATM:g} atm, 6 cells", flush=True)
    step = api.ChemistryStep(mech, method="team", jacobian="analytic", pressure=P)
    try:
        f, Ja = step.equations(y)
    finally:
        step.close()
    fo = _rhs(oracle, mech, y, P)
    scale = np.abs(fo[:mech.ns]).max(axis=0, keepdims=True)
    err = np.abs(f[:mech.ns] - fo[:mech.ns]) / scale
    errT = np.abs(f[mech.ns] - fo[mech.ns]) / np.abs(fo[mech.ns])
    assert err.max() < 1e-11 and errT.max() < 1e-11, (err.max(), errT.max())
    worst = 0.0
    for c in range(y.shape[1]):
        Jc = _cd_jacobian(oracle, mech, y[:, c], P)
        e = np.abs(Ja[c] - Jc) / (np.abs(Jc).max(axis=1, keepdims=True) + 1e-300)
        worst = max(worst, float(e.max()))
        assert e.max() < 1e-5, (c, np.unravel_index(e.argmax(), e.shape), e.max())
    print(f"P = {P / ATM:g} atm: rhs err {max(err.max(), errT.max()):.3g}, jacobian err {worst:.3g}", flush=True)


CASES = [(P, m, j) for P in PRESSURES[:2] for m, j in (("team", "numerical"), ("plain", "numerical"),
                                                       ("wide", "numerical"), ("team", "analytic"),
                                                       ("wide", "analytic"))]
CASES += [(P, "team", "numerical") for P in PRESSURES[2:]]


@pytest.mark.parametrize("P,method,jac", CASES)
def test_step(hip_device, oracle, mech, P, method, jac):
    y = premixed_cells(mech, 48, seed=5)
    print(f"step P = {P / ATM:g} atm method={method} jacobian={jac} cells=48 dt={DT}", flush=True)
    kw = dict(rtol=RTOL, atol=ATOL, pressure=P)
    yo, st_o, *_ = oracle.chemistry_step(mech, y, DT, threads=8, **kw)
    ode = api.ODE_Parameters()
    ode.SetRelativeTolerance(RTOL)
    ode.SetAbsoluteTolerance(ATOL)
    step = api.ChemistryStep(mech, ode, method=method, jacobian=jac, pressure=P)
    try:
        yg, st_g, q_g, c_g = step.integrate(y, DT)
    finally:
        step.close()
    ignited = int((yo[mech.ns] - y[mech.ns] > 100.0).sum())
    variant = method + ("-analytic" if jac == "analytic" else "")
    env = oracle_envelope(oracle, mech, y, DT, RTOL, ATOL, yo, pressure=P)
    if newton_matrix_changed(variant):
        env = np.maximum(env, oracle_jacobian_envelope(oracle, mech, y, DT, RTOL, ATOL, yo, pressure=P))
        yo, st_o = reconcile_statuses(oracle, mech, y, DT, RTOL, ATOL, st_g, st_o, yo, what=f"P={P}/{variant}",
                                      pressure=P)
    assert np.array_equal(st_g, st_o), (st_g, st_o)
    (cell, _), env = assert_parity_adaptive(oracle, mech, y, DT, RTOL, ATOL, yo, yg, env, what=f"P={P}/{variant}",
                                            pressure=P)
    assert_element_conservation(mech, y, yg, c_g[:, 10], what=f"P={P}/{method}/{jac}", y_ref=yo)
    print(f"P = {P / ATM:g} atm {method}/{jac}: max ratio {cell.max():.3g}; {ignited} of 48 ignite", flush=True)


@pytest.mark.parametrize("method", ["team", "wide", "plain"])
def test_constant_volume(hip_device, oracle, mech, method):
    """Constant volume: the PLOG constants follow the pressure of the state."""
    y = premixed_cells(mech, 32, seed=9)
    rho = 4.0 * ATM * 28.0 / (8314.4621 * y[mech.ns])          # about 4 atm at the initial temperature
    print(f"constant volume, method={method}, 32 cells", flush=True)
    kw = dict(rtol=RTOL, atol=ATOL, constant_volume=True)
    yo, st_o, *_ = oracle.chemistry_step(mech, y, DT, threads=8, rho=rho, **kw)
    env = oracle_envelope(oracle, mech, y, DT, RTOL, ATOL, yo, rho=rho, constant_volume=True)
    ode = api.ODE_Parameters()
    ode.SetRelativeTolerance(RTOL)
    ode.SetAbsoluteTolerance(ATOL)
    step = api.ChemistryStep(mech, ode, method=method, constant_volume=True)
    try:
        yg, st_g, *_ = step.integrate(y, DT, rho=rho)
    finally:
        step.close()
    assert np.array_equal(st_g, st_o)
    cell, _ = assert_parity(yg, yo, RTOL, ATOL, env, what=f"cv/{method}")
    print(f"constant volume {method}: max ratio {cell.max():.3g}", flush=True)


def _device_batch(y, dt):
    import torch
    dev = torch.device("cuda:0")
    n = y.shape[1]
    b = dict(n=n, y=torch.from_numpy(y).to(dev), dt=torch.full((n,), dt, dtype=torch.float64, device=dev))
    b["out"] = torch.empty_like(b["y"])
    b["st"] = torch.zeros(n, dtype=torch.int32, device=dev)
    b["q"] = torch.zeros(n, dtype=torch.float64, device=dev)
    b["cnt"] = torch.zeros((n, 12), dtype=torch.int32, device=dev)
    return b


def _run(step, b, stream, order=None):
    step.integrate_device(b["n"], b["dt"].data_ptr(), b["y"].data_ptr(), b["out"].data_ptr(), b["st"].data_ptr(),
                          b["q"].data_ptr(), b["cnt"].data_ptr(), stream.cuda_stream,
                          order.data_ptr() if order is not None else None)


def _result(b):
    import torch
    torch.cuda.synchronize()
    return b["out"].cpu().numpy().copy(), b["st"].cpu().numpy().copy(), b["cnt"].cpu().numpy().copy()


@pytest.mark.parametrize("method,jac", [("team", "numerical"), ("team", "analytic"), ("wide", "analytic")])
def test_order_does_not_change_results(hip_device, mech, method, jac):
    """Cost order and its reverse give the unordered call's results bit for bit."""
    import torch
    y = premixed_cells(mech, 256, seed=21)
    print(f"order test {method}/{jac}, forms mechanism, 256 cells", flush=True)
    ode = api.ODE_Parameters()
    ode.SetRelativeTolerance(RTOL)
    ode.SetAbsoluteTolerance(ATOL)
    step = api.ChemistryStep(mech, ode, method=method, jacobian=jac, pressure=3.7 * ATM)
    b = _device_batch(y, DT)
    s = torch.cuda.current_stream()
    try:
        _run(step, b, s)
        y0, st0, c0 = _result(b)
        order = torch.empty(b["n"], dtype=torch.int32, device="cuda:0")
        step.cost_order(b["n"], b["cnt"].data_ptr(), order.data_ptr(), s.cuda_stream)
        torch.cuda.synchronize()
        assert np.array_equal(np.sort(order.cpu().numpy()), np.arange(b["n"]))
        _run(step, b, s, order)
        y1, st1, c1 = _result(b)
        _run(step, b, s, torch.flip(order, [0]).contiguous())
        y2, st2, c2 = _result(b)
    finally:
        step.close()
    for ya, sa, ca in ((y1, st1, c1), (y2, st2, c2)):
        assert np.array_equal(ya, y0) and np.array_equal(sa, st0) and np.array_equal(ca, c0)
    assert np.all(st0 >= 0)
    print(f"{method}/{jac}: ordered results identical; steps {c0[:, 0].min()}-{c0[:, 0].max()}", flush=True)


def test_graph_replay_matches_eager(hip_device, mech):
    """A team step of the forms mechanism captured in a HIP graph: three replays, with
    eager steps between them, each equal to the eager result."""
    import torch
    y = premixed_cells(mech, 512, seed=23)
    print("graph test, forms mechanism, 512 cells", flush=True)
    ode = api.ODE_Parameters()
    ode.SetRelativeTolerance(RTOL)
    ode.SetAbsoluteTolerance(ATOL)
    step = api.ChemistryStep(mech, ode, method="team", pressure=0.02 * ATM)
    b = _device_batch(y, DT)
    s = torch.cuda.current_stream()
    try:
        _run(step, b, s)
        y0, st0, c0 = _result(b)
        side = torch.cuda.Stream()
        side.wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(side):
            with torch.cuda.graph(g, stream=side):
                _run(step, b, side)
        s.wait_stream(side)
        torch.cuda.synchronize()
        for rep in range(3):
            b["out"].zero_()
            b["st"].fill_(-99)
            g.replay()
            y1, st1, c1 = _result(b)
            assert np.array_equal(y1, y0) and np.array_equal(st1, st0) and np.array_equal(c1, c0), rep
            _run(step, b, s)
            y2, _, _ = _result(b)
            assert np.array_equal(y2, y0)
    finally:
        step.close()
    print("graph: 3 replays identical to eager", flush=True)
