"""GPU: the transport part of the per-cell property refresh (lsm_transport_properties)
against the oracle on the 14-, 53- and 160-species workload fields and the
reference's CH4 skeletal mechanism, with every viscosity model; the pure-species
limit; the solver's options (simplified properties, Lewis-number diffusivities, soot
corrections, Soret coefficients) against the oracle's restatement of properties.H; and
the bandwidth of the kernel on a 262,144-cell field.

Tolerances: mu, lambda and the diffusion coefficients to 1e-12 relative; the thermal
diffusion ratios theta_k = x_k sum_j theta_kj(T) x_j (and Dsoret) to 1e-12 of the sum of
the absolute monomials, since the fitted cubics and the sum over partners can cancel far
below their terms.  The kernel
sums over species in a different order from the oracle (wave reductions), so results
agree to rounding, not bit for bit.
"""
import os
import time

import numpy as np
import pytest

from laminarsmoke_amd import api
from laminarsmoke_amd.chemkin import load_chemkin
from laminarsmoke_amd.fields import workload_cells
from laminarsmoke_amd.mechanism import workload_mechanism
from laminarsmoke_amd.transport import (VISCOSITY_HERNING, VISCOSITY_MATHUR_SAXENA, VISCOSITY_WILKE, Transport,
                                        TransportOptions, synthetic_transport)

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _oracle(oracle, mech, tr, y, p):
    n = y.shape[1]
    ns = mech.ns
    mu, lam = np.empty(n), np.empty(n)
    dmix, teta = np.empty((ns, n)), np.empty((ns, n))
    mw = mech.mw
    for c in range(n):
        u = np.sum(y[:ns, c] / mw)
        x = ((1.0 / u) * y[:ns, c]) / mw    # x_k = MW Y_k / MW_k, MW = 1 / sum(Y/MW)
        mu[c], lam[c], dmix[:, c], teta[:, c] = oracle.transport(tr, y[ns, c], p[c], x)
    return mu, lam, dmix, teta


def _teta_terms(mech, tr, y):
    """x_k sum_j (|a0| + |a1 T| + |a2 T^2| + |a3 T^3|)_kj x_j per species and cell: the size
    of the terms whose signed sums are theta_kj(T) and theta_k, which bounds their rounding
    error (the cubic in T and the sum over partners can cancel far below their terms, and
    the kernel's fused multiply-adds round differently from the oracle's separate ones)."""
    ns = mech.ns
    MW = 1.0 / np.sum(y[:ns] / mech.mw[:, None], axis=0)
    x = MW * y[:ns] / mech.mw[:, None]
    T = y[ns]
    out = np.zeros((ns, y.shape[1]))
    for q, k in enumerate(tr.teta_sp):
        f = tr.teta_fit[q]
        th = np.abs(f[:, :1]) + np.abs(f[:, 1:2] * T) + np.abs(f[:, 2:3] * T ** 2) + np.abs(f[:, 3:4] * T ** 3)
        out[k] = x[k] * np.sum(th * x, axis=0)
    return out, MW


def _assert_teta(got, want, bound, what):
    r = np.abs(got - want) / (bound + 1e-300)
    k, c = np.unravel_index(np.argmax(r), r.shape)
    assert r[k, c] <= 1.0, (f"{what}: species {k} cell {c}: got {got[k, c]!r} want {want[k, c]!r} "
                            f"bound {bound[k, c]!r}")


def _check(mech, tr, y, p, oracle, what):
    step = api.ChemistryStep(mech)
    try:
        step.set_transport(tr)
        mu, lam, dmix, teta = step.transport(y, p)
    finally:
        step.close()
    mo, lo, do, to = _oracle(oracle, mech, tr, y, p)
    np.testing.assert_allclose(mu, mo, rtol=1e-12, err_msg=f"{what}: mu")
    np.testing.assert_allclose(lam, lo, rtol=1e-12, err_msg=f"{what}: lambda")
    np.testing.assert_allclose(dmix, do, rtol=1e-12, err_msg=f"{what}: dmix")
    terms, _ = _teta_terms(mech, tr, y)
    _assert_teta(teta, to, 1e-12 * terms, f"{what}: teta")
    print(f"{what}: mu {mu.min():.3g}-{mu.max():.3g} kg/m/s, lambda {lam.min():.3g}-{lam.max():.3g} W/m/K, "
          f"max rel dev mu {np.max(np.abs(mu / mo - 1)):.2g}, dmix {np.max(np.abs(dmix / do - 1)):.2g}", flush=True)


@pytest.mark.parametrize("cfg", [1, 2, 4])
@pytest.mark.parametrize("model", [VISCOSITY_WILKE, VISCOSITY_HERNING, VISCOSITY_MATHUR_SAXENA])
def test_workload_fields(hip_device, oracle, cfg, model):
    mech, y = workload_cells(cfg, n=96)
    tr = synthetic_transport(mech, seed=cfg, model=model)
    p = np.random.default_rng(cfg).uniform(2e4, 2e6, y.shape[1])
    print(f"transport cfg{cfg} model {model}: 96 cells, {mech.ns} species", flush=True)
    _check(mech, tr, y, p, oracle, f"cfg{cfg}/model{model}")


def test_ch4_skeletal(hip_device, oracle):
    mech = load_chemkin(os.path.join(GOLDEN, "POLIMI_CH4_SKELETAL_1412.CKI"),
                        os.path.join(GOLDEN, "POLIMI_CH4_SKELETAL_1412_thermo.CKT"))
    from laminarsmoke_amd.fields import premixed_cells
    y = premixed_cells(mech, 64, seed=4)
    tr = synthetic_transport(mech, seed=11)
    print("transport CH4 skeletal: 64 cells", flush=True)
    _check(mech, tr, y, np.full(64, 101325.0), oracle, "ch4")


def test_pure_species(hip_device):
    """A cell of one species returns that species' fitted mu and lambda."""
    mech = workload_mechanism(14)
    tr = synthetic_transport(mech, seed=2)
    n = mech.ns
    y = np.zeros((mech.ns + 1, n))
    y[np.arange(n), np.arange(n)] = 1.0
    y[mech.ns] = 1234.0
    step = api.ChemistryStep(mech)
    try:
        step.set_transport(tr)
        mu, lam, _, _ = step.transport(y, species=False)
    finally:
        step.close()
    lt = np.log(1234.0)
    v = np.array([1.0, lt, lt * lt, lt ** 3])
    np.testing.assert_allclose(mu, np.exp(tr.eta_fit @ v), rtol=1e-13)
    np.testing.assert_allclose(lam, np.exp(tr.lambda_fit @ v), rtol=1e-13)


def test_bandwidth(hip_device):
    """262,144 cells of the 53-species field, device pointers: time per field refresh."""
    import torch
    mech, y = workload_cells(2, n=262144)
    tr = synthetic_transport(mech, seed=2)
    dev = torch.device("cuda:0")
    n = y.shape[1]
    yd = torch.from_numpy(y).to(dev)
    mu, lam = torch.empty(n, dtype=torch.float64, device=dev), torch.empty(n, dtype=torch.float64, device=dev)
    dmix = torch.empty((mech.ns, n), dtype=torch.float64, device=dev)
    step = api.ChemistryStep(mech)
    try:
        step.set_transport(tr)
        for _ in range(2):
            step.transport_device(n, yd.data_ptr(), mu.data_ptr(), lam.data_ptr(), dmix_ptr=dmix.data_ptr())
        torch.cuda.synchronize()
        t = time.time()
        reps = 5
        for _ in range(reps):
            step.transport_device(n, yd.data_ptr(), mu.data_ptr(), lam.data_ptr(), dmix_ptr=dmix.data_ptr())
        torch.cuda.synchronize()
        dt = (time.time() - t) / reps
    finally:
        step.close()
    assert torch.isfinite(mu).all() and torch.isfinite(dmix).all()
    nbytes = 8.0 * n * ((mech.ns + 1) + 2 + mech.ns)
    print(f"transport 262144 cells, {mech.ns} species: {dt * 1e3:.3f} ms per field, {n / dt:.3g} cells/s, "
          f"{nbytes / dt / 1e9:.0f} GB/s of state in + mu, lambda, dmix out", flush=True)


def test_edge_cases(hip_device):
    """Empty batch; p = None is the handle's pressure; errors before lsm_set_transport and
    for fits of another species count; replacing the fits takes effect."""
    mech, y = workload_cells(1, n=8)
    tr = synthetic_transport(mech, seed=5)
    step = api.ChemistryStep(mech, pressure=2.5e5)
    try:
        with pytest.raises(RuntimeError, match="lsm_set_transport"):
            step.transport(y)
        other = synthetic_transport(workload_mechanism(53), seed=1)
        with pytest.raises(RuntimeError, match="ns differs"):
            step.set_transport(other)
        step.set_transport(tr)
        mu0, lam0, d0, t0 = step.transport(y[:, :0])
        assert mu0.size == 0 and d0.shape == (mech.ns, 0)
        mu1, lam1, d1, _ = step.transport(y)
        mu2, lam2, d2, _ = step.transport(y, np.full(8, 2.5e5))
        assert np.array_equal(mu1, mu2) and np.array_equal(d1, d2)
        mu3, _, d3, _ = step.transport(y, np.full(8, 5e5))
        np.testing.assert_allclose(d3 * 2.0, d1, rtol=1e-13)           # D ~ 1/P
        assert np.array_equal(mu3, mu1)                                 # mu does not depend on P
        tr2 = Transport(tr.mw, tr.lambda_fit, tr.eta_fit + np.array([np.log(2.0), 0, 0, 0]), tr.gamma_fit,
                        tr.teta_sp, tr.teta_fit)
        step.set_transport(tr2)
        mu4, _, _, _ = step.transport(y)
        np.testing.assert_allclose(mu4, 2.0 * mu1, rtol=1e-13)        # every eta doubled
    finally:
        step.close()


def _options(kind, ns, seed):
    rng = np.random.default_rng(seed)
    le = rng.uniform(0.3, 2.0, ns)
    soot = dict(soot_reference=3, soot_species=[1, 3, ns - 2], soot_correction=[0.5, 2.0, 0.25])
    return {"simplified": TransportOptions(simplified=True),
            "lewis": TransportOptions(diffusivity="lewis-numbers", lewis=le),
            "lewis_all": TransportOptions(diffusivity="lewis-numbers", lewis=le, simplified=True, soret=True, **soot),
            "multi_soret_soot": TransportOptions(soret=True, **soot)}[kind]


@pytest.mark.parametrize("cfg", [1, 2])
@pytest.mark.parametrize("kind", ["simplified", "lewis", "lewis_all", "multi_soret_soot"])
def test_refresh_options(hip_device, oracle, cfg, kind):
    mech, y = workload_cells(cfg, n=64)
    tr = synthetic_transport(mech, seed=cfg + 7)
    opts = _options(kind, mech.ns, cfg)
    p = np.random.default_rng(cfg + 1).uniform(5e4, 1e6, y.shape[1])
    print(f"refresh options {kind}, cfg{cfg}: 64 cells, {mech.ns} species", flush=True)
    step = api.ChemistryStep(mech)
    try:
        step.set_transport(tr)
        step.set_transport_options(opts)
        mu, lam, dmix, teta = step.transport(y, p)
        step.set_transport_options(None)
        mu0, _, d0, _ = step.transport(y, p)
    finally:
        step.close()
    want = [oracle.refresh_cell(mech, tr, opts, p[c], y[:, c]) for c in range(y.shape[1])]
    mo = np.array([w[0] for w in want])
    lo = np.array([w[1] for w in want])
    do = np.stack([w[2] for w in want], axis=1)
    to = np.stack([w[3] for w in want], axis=1)
    np.testing.assert_allclose(mu, mo, rtol=1e-12, err_msg=f"{kind}: mu")
    np.testing.assert_allclose(lam, lo, rtol=1e-12, err_msg=f"{kind}: lambda")
    np.testing.assert_allclose(dmix, do, rtol=1e-12, err_msg=f"{kind}: dmix")
    terms, MW = _teta_terms(mech, tr, y)
    if opts.soret:                      # Dsoret_k = D_k theta_k MW_k / MW
        terms = terms * np.abs(do) * mech.mw[:, None] / MW
    _assert_teta(teta, to, 1e-12 * terms, f"{kind}: teta / Dsoret")
    base = [oracle.refresh_cell(mech, tr, None, p[c], y[:, c]) for c in range(y.shape[1])]
    np.testing.assert_allclose(mu0, [b[0] for b in base], rtol=1e-12, err_msg="defaults restored: mu")
    np.testing.assert_allclose(d0, np.stack([b[2] for b in base], axis=1), rtol=1e-12, err_msg="defaults: dmix")
    print(f"{kind} cfg{cfg}: mu {mu.min():.3g}-{mu.max():.3g}, D {dmix.min():.3g}-{dmix.max():.3g} m2/s", flush=True)


def test_refresh_soret_without_dmix(hip_device, oracle):
    """Dsoret needs the diffusion coefficients even when the caller does not ask for them."""
    import torch
    mech, y = workload_cells(1, n=32)
    tr = synthetic_transport(mech, seed=3)
    dev = torch.device("cuda:0")
    yd = torch.from_numpy(y).to(dev)
    n = y.shape[1]
    mu, lam = torch.empty(n, dtype=torch.float64, device=dev), torch.empty(n, dtype=torch.float64, device=dev)
    ds = torch.empty((mech.ns, n), dtype=torch.float64, device=dev)
    step = api.ChemistryStep(mech)
    try:
        step.set_transport(tr)
        step.set_transport_options(TransportOptions(soret=True))
        step.transport_device(n, yd.data_ptr(), mu.data_ptr(), lam.data_ptr(), teta_ptr=ds.data_ptr())
        torch.cuda.synchronize()
    finally:
        step.close()
    want = np.stack([oracle.refresh_cell(mech, tr, TransportOptions(soret=True), 101325.0, y[:, c])[3]
                     for c in range(n)], axis=1)
    got = ds.cpu().numpy()
    terms, MW = _teta_terms(mech, tr, y)
    d = np.stack([oracle.refresh_cell(mech, tr, None, 101325.0, y[:, c])[2] for c in range(n)], axis=1)
    _assert_teta(got, want, 1e-12 * terms * d * mech.mw[:, None] / MW, "Dsoret")


def test_refresh_options_rejected(hip_device):
    mech, _ = workload_cells(1, n=8)
    step = api.ChemistryStep(mech)
    try:
        bad = [TransportOptions(diffusivity="lewis-numbers"),
               TransportOptions(diffusivity="lewis-numbers", lewis=np.zeros(mech.ns)),
               TransportOptions(soot_reference=mech.ns, soot_species=[0], soot_correction=[1.0]),
               TransportOptions(soot_reference=0, soot_species=[1, 1], soot_correction=[1.0, 1.0]),
               TransportOptions(soot_species=[1], soot_correction=[1.0])]
        for o in bad:
            with pytest.raises(api.LsmError):
                step.set_transport_options(o)
        step.set_transport_options(TransportOptions(simplified=True))
        step.set_transport_options(None)
    finally:
        step.close()
