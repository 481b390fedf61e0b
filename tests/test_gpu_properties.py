"""GPU: the thermodynamic part of the per-cell property refresh (lsm_properties,
solvers/laminarSMOKE/properties.H:41-131) against the oracle's NASA-7 species
properties combined by the definitions of the mixture quantities:
  MW = 1 / sum(Y_k / MW_k), X_k = MW Y_k / MW_k, cTot = p / (R T), psi = cTot MW / p,
  h = sum X_k H_k / MW, cp = sum X_k Cp_k / MW, cv = (sum X_k Cp_k - R) / MW,
  CpSpecies_k = Cp_k / MW_k;
on the 53-species coflow field (handle pressure, and a per-cell pressure array) and the
reference's H2/CO mechanism.  The last test times the kernel on the 262,144-cell field
and reports its HBM bandwidth (the kernel is bandwidth-bound: ~1.3 KB per cell for
53 species against ~1 kFLOP).
"""
import os

import numpy as np
import pytest
import torch

from laminarsmoke_amd import api
from laminarsmoke_amd.chemkin import load_chemkin
from laminarsmoke_amd.fields import coflow_field, premixed_cells
from laminarsmoke_amd.mechanism import workload_mechanism

pytestmark = pytest.mark.gpu

R = 8314.4621
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def expected(oracle, mech, y, p):
    n = y.shape[1]
    out = np.empty((6, n))
    x = np.empty((mech.ns, n))
    cpk = np.empty((mech.ns, n))
    for c in range(n):
        T = y[-1, c]
        cp, h, _ = oracle.species_thermo(mech, T)
        MW = 1.0 / np.sum(y[:-1, c] / mech.mw)
        xc = MW * y[:-1, c] / mech.mw
        cTot = p[c] / R / T
        out[:, c] = (MW, cTot, cTot * MW / p[c], np.sum(xc * h) * R * T / MW, np.sum(xc * cp) * R / MW,
                     (np.sum(xc * cp) * R - R) / MW)
        x[:, c] = xc
        cpk[:, c] = cp * R / mech.mw
    return out, x, cpk


def cases():
    m53 = workload_mechanism(53)
    y53, _ = coflow_field(m53, 64, 64, 22)
    mh = load_chemkin(os.path.join(GOLDEN, "POLIMI_H2CO_1412.CKI"), os.path.join(GOLDEN, "POLIMI_H2CO_1412_thermo.CKT"))
    return [("cfg2", m53, y53[:, ::8]), ("h2co", mh, premixed_cells(mh, 256, seed=4, T_lo=300.0, T_hi=2500.0))]


@pytest.mark.parametrize("k", [0, 1])
@pytest.mark.parametrize("per_cell_p", [False, True])
def test_properties(hip_device, oracle, k, per_cell_p):
    name, mech, y = cases()[k]
    n = y.shape[1]
    p = np.linspace(0.8e5, 3e5, n) if per_cell_p else np.full(n, 101325.0)
    print(f"properties {name} cells={n} per-cell pressure={per_cell_p}", flush=True)
    st = api.ChemistryStep(mech)
    try:
        out, x, cpk = st.properties(y, p if per_cell_p else None)
    finally:
        st.close()
    eo, ex, ec = expected(oracle, mech, y, p)
    for i, nm in enumerate(api.ChemistryStep.PROPS):
        np.testing.assert_allclose(out[i], eo[i], rtol=1e-13, atol=1e-13 * np.abs(eo[i]).max(), err_msg=nm)
    np.testing.assert_allclose(x, ex, rtol=1e-14, atol=1e-16)
    np.testing.assert_allclose(cpk, ec, rtol=1e-13)
    np.testing.assert_allclose(x.sum(0), 1.0, atol=1e-14)


def test_properties_bandwidth(hip_device):
    mech = workload_mechanism(53)
    y, _ = coflow_field(mech, 512, 512, 22)
    n = y.shape[1]
    dev = torch.device("cuda:0")
    yd = torch.from_numpy(y).to(dev)
    out = torch.empty((6, n), dtype=torch.float64, device=dev)
    x = torch.empty((mech.ns, n), dtype=torch.float64, device=dev)
    cpk = torch.empty((mech.ns, n), dtype=torch.float64, device=dev)
    st = api.ChemistryStep(mech)
    s = torch.cuda.current_stream()
    try:
        run = lambda: st.properties_device(n, yd.data_ptr(), out.data_ptr(), None, x.data_ptr(), cpk.data_ptr(),
                                           s.cuda_stream)
        run()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 20
        a.record(s)
        for _ in range(reps):
            run()
        b.record(s)
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / reps
        o = out.cpu().numpy()
    finally:
        st.close()
    assert np.all(np.isfinite(o)) and np.allclose(x.sum(0).cpu().numpy(), 1.0, atol=1e-13)
    nbytes = 8.0 * n * ((mech.ns + 1) + 6 + 2 * mech.ns)      # state in; props, X, Cp_k out
    print(f"properties 262144 cells, 53 species: {ms:.3f} ms per field, {n / ms * 1e3:.3g} cells/s, "
          f"{nbytes / ms / 1e6:.0f} GB/s of algorithmic traffic (HBM peak 8000)", flush=True)
