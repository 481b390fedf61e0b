"""Pins of the oracle's transport properties and of the Transport XML block (CPU only).

The mixture rules are checked against the textbook forms they come from, written
here independently of the map's pair-table arithmetic:
* Wilke's viscosity (Reid, Prausnitz, Poling, The properties of gases and liquids,
  eq. 9-5.13/14): eta = sum_i x_i eta_i / sum_j x_j phi_ij,
  phi_ij = [1 + (eta_i/eta_j)^1/2 (M_j/M_i)^1/4]^2 / [8 (1 + M_i/M_j)]^1/2;
* a pure species returns its own eta and lambda under every mixture rule;
* the mixture-averaged diffusion coefficient of a binary mixture is
  D_12 M_2 / MWmix (from (1 - y_1) / (x_2 / D_12)), and D scales as 1/P;
* constant thermal diffusion ratios theta give theta x_i.
"""
import numpy as np
import pytest

from laminarsmoke_amd.mechanism import workload_mechanism
from laminarsmoke_amd.transport import (VISCOSITY_HERNING, VISCOSITY_MATHUR_SAXENA, VISCOSITY_WILKE, Transport,
                                        read_transport_xml, synthetic_transport, write_transport_block)


def _species(tr, T):
    lt = np.log(T)
    v = np.array([1.0, lt, lt * lt, lt ** 3])
    return np.exp(tr.lambda_fit @ v), np.exp(tr.eta_fit @ v)


def _pair_D(tr, j, k, T, P):
    j, k = min(j, k), max(j, k)
    ns = tr.ns
    q = j * ns - j * (j + 1) // 2 + (k - j - 1)
    lt = np.log(T)
    return np.exp(tr.gamma_fit[q] @ np.array([1.0, lt, lt * lt, lt ** 3])) / (P / 1e5)


@pytest.fixture(scope="module")
def tr14():
    return synthetic_transport(workload_mechanism(14), seed=3)


@pytest.mark.parametrize("T", [320.0, 1200.0, 2900.0])
def test_wilke_against_textbook(oracle, tr14, T):
    tr = tr14
    lam, eta = _species(tr, T)
    x = np.random.default_rng(int(T)).dirichlet(np.full(tr.ns, 0.7))
    M = tr.mw
    phi = (1 + np.sqrt(eta[:, None] / eta[None, :]) * (M[None, :] / M[:, None]) ** 0.25) ** 2 / \
        np.sqrt(8 * (1 + M[:, None] / M[None, :]))
    mu_ref = np.sum(x * eta / (phi @ x))
    mu, lmix, _, _ = oracle.transport(tr, T, 101325.0, x)
    assert mu == pytest.approx(mu_ref, rel=1e-13)
    assert lmix == pytest.approx(0.5 * (x @ lam + 1.0 / (x @ (1.0 / lam))), rel=1e-13)


@pytest.mark.parametrize("model", [VISCOSITY_WILKE, VISCOSITY_HERNING, VISCOSITY_MATHUR_SAXENA])
def test_pure_species(oracle, tr14, model):
    tr = Transport(tr14.mw, tr14.lambda_fit, tr14.eta_fit, tr14.gamma_fit, tr14.teta_sp, tr14.teta_fit, model)
    T = 1500.0
    lam, eta = _species(tr, T)
    for k in (0, 5, 13):
        x = np.zeros(tr.ns)
        x[k] = 1.0
        mu, lmix, _, _ = oracle.transport(tr, T, 101325.0, x)
        assert mu == pytest.approx(eta[k], rel=1e-14)
        assert lmix == pytest.approx(lam[k], rel=1e-14)


def test_herning_and_mathur_saxena(oracle, tr14):
    T = 900.0
    lam, eta = _species(tr14, T)
    x = np.random.default_rng(5).dirichlet(np.ones(tr14.ns))
    sq = np.sqrt(tr14.mw)
    for model, expect in ((VISCOSITY_HERNING, (x * eta * sq).sum() / (x * sq).sum()),
                          (VISCOSITY_MATHUR_SAXENA, 0.5 * (x @ eta + 1.0 / (x @ (1.0 / eta))))):
        tr = Transport(tr14.mw, tr14.lambda_fit, tr14.eta_fit, tr14.gamma_fit, tr14.teta_sp, tr14.teta_fit, model)
        assert oracle.transport(tr, T, 101325.0, x)[0] == pytest.approx(expect, rel=1e-13)


@pytest.mark.parametrize("P", [5e3, 101325.0, 3e6])
def test_binary_diffusion(oracle, tr14, P):
    """Two species present: D_1m = D_12 M_2 / MWmix (to the map's 1e-14 mole-fraction floor)."""
    T = 1400.0
    x = np.zeros(tr14.ns)
    x[2], x[7] = 0.3, 0.7
    _, _, dmix, _ = oracle.transport(tr14, T, P, x)
    D27 = _pair_D(tr14, 2, 7, T, P)
    M = tr14.mw
    mwmix = 0.3 * M[2] + 0.7 * M[7]
    assert dmix[2] == pytest.approx(D27 * M[7] / mwmix, rel=1e-9)
    assert dmix[7] == pytest.approx(D27 * M[2] / mwmix, rel=1e-9)
    d1 = oracle.transport(tr14, T, 101325.0, x)[2]
    np.testing.assert_allclose(dmix * P, d1 * 101325.0, rtol=1e-13)


def test_mixture_averaged_diffusion(oracle, tr14):
    """D_km = (1 - y_k) / sum_{j != k} x_j / D_kj with the map's floored mole fractions."""
    T, P = 1800.0, 2e5
    x = np.random.default_rng(9).dirichlet(np.full(tr14.ns, 0.5))
    xc = (x + 1e-14) / (1 + 1e-14 * tr14.ns)
    M = tr14.mw
    mw = xc @ M
    _, _, dmix, _ = oracle.transport(tr14, T, P, x)
    for k in range(tr14.ns):
        s = sum(xc[j] / _pair_D(tr14, k, j, T, P) for j in range(tr14.ns) if j != k)
        assert dmix[k] == pytest.approx((1 - xc[k] * M[k] / mw) / s, rel=1e-12)


def test_thermal_diffusion_ratios(oracle, tr14):
    """Constant theta_ij = c_i: teta_i = c_i x_i sum_j x_j = c_i x_i; other species 0."""
    ns = tr14.ns
    teta_fit = np.zeros_like(tr14.teta_fit)
    c = np.arange(1, len(tr14.teta_sp) + 1) * 0.05
    teta_fit[:, :, 0] = c[:, None]
    tr = Transport(tr14.mw, tr14.lambda_fit, tr14.eta_fit, tr14.gamma_fit, tr14.teta_sp, teta_fit)
    x = np.random.default_rng(2).dirichlet(np.ones(ns))
    teta = oracle.transport(tr, 1000.0, 101325.0, x)[3]
    expect = np.zeros(ns)
    expect[tr14.teta_sp] = c * x[tr14.teta_sp]
    np.testing.assert_allclose(teta, expect, rtol=1e-14, atol=1e-300)
    assert set(tr14.teta_sp.tolist()) == {k for k in range(ns) if tr14.mw[k] <= 20.0}


def test_transport_block_round_trip(tr14):
    m = workload_mechanism(14)
    for model in (VISCOSITY_WILKE, VISCOSITY_MATHUR_SAXENA):
        tr = Transport(tr14.mw, tr14.lambda_fit, tr14.eta_fit, tr14.gamma_fit, tr14.teta_sp, tr14.teta_fit, model)
        doc = f"<opensmoke><NumberOfSpecies> {m.ns} </NumberOfSpecies>\n{write_transport_block(tr)}\n</opensmoke>"
        back = read_transport_xml(doc)
        for f in ("mw", "lambda_fit", "eta_fit", "gamma_fit", "teta_sp", "teta_fit"):
            assert np.array_equal(getattr(back, f), getattr(tr, f)), f
        assert back.viscosity_model == model


def test_hand_written_block():
    """Two species, H2 (MW 2.016, with thermal diffusion ratios) and N2: the order of
    ImportCoefficientsFromASCIIFile (count, then per species MW, 4 lambda, 4 eta, the
    pairs with the species after it, the theta rows of a light species, then E)."""
    doc = """<opensmoke><NumberOfSpecies> 2 </NumberOfSpecies>
<Transport type="CHEMKIN">
1
2.016 1 2 3 4 5 6 7 8 9 10 11 12 0.1 0.2 0.3 0.4 0.5 0.6 0.7 0.8
28.014 13 14 15 16 17 18 19 20
E
</Transport>
<ViscosityModel> Herning </ViscosityModel>
</opensmoke>"""
    tr = read_transport_xml(doc)
    assert tr.mw.tolist() == [2.016, 28.014]
    assert tr.lambda_fit.tolist() == [[1, 2, 3, 4], [13, 14, 15, 16]]
    assert tr.eta_fit.tolist() == [[5, 6, 7, 8], [17, 18, 19, 20]]
    assert tr.gamma_fit.tolist() == [[9, 10, 11, 12]]
    assert tr.teta_sp.tolist() == [0] and tr.teta_fit.tolist() == [[[0.1, 0.2, 0.3, 0.4], [0.5, 0.6, 0.7, 0.8]]]
    assert tr.viscosity_model == VISCOSITY_HERNING


def test_synthetic_fits_are_physical(tr14):
    """Synthetic fits give gas-like magnitudes: eta ~ 1e-5-1e-4 kg/m/s, D ~ 1e-5-1e-3 m2/s."""
    lam, eta = _species(tr14, 1000.0)
    assert np.all((eta > 5e-6) & (eta < 3e-4)) and np.all((lam > 1e-3) & (lam < 2.0))
    assert 1e-6 < _pair_D(tr14, 0, 1, 300.0, 101325.0) < 1e-3
