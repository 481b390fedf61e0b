"""GPU: the 1,048,576-cell stiffness-spread workload (53 species / 325 reactions, T 300-2500 K,
radicals log-uniform 1e-12 - 1e-3, dt 1e-5 s, rtol 1e-7, atol 1e-12), the benchmark's default.

* All cells in one call of the team kernel with the analytic Jacobian (the benchmark's
  configuration), checked by properties that hold cell by cell at any size: every final state
  finite, T within [yMin, yMax] = [200, 6000] K (MethodGear.hpp:539-569), Y_k >= 0 and
  sum Y = 1 to 1e-14 (normalizeMassFractions, chemistry_DI.H:219), element mass conserved
  (tests/parity.py's bound, scaled by each cell's clamp events), statuses 2 or a failure status
  of OdeSolverUtilities.h:83-99, counters consistent (a successful cell took at least one step).
* Against the oracle, 64 cells stratified by temperature (8 in each of 8 bands of 275 K), the 8
  costliest cells by the full call's step counters and up to 8 cells where it failed: identical
  statuses and the per-cell parity bound of tests/parity.py, for the full call's results
  (analytic Jacobian) and for a separate call of the team kernel with the numerical Jacobian.
  The analytic Jacobian changes the Newton matrix on purpose, so a status that differs from the
  oracle's must be the oracle's own with another valid FD Jacobian (tests/parity.py
  reconcile_statuses): on cell 212156 both return -6 after 1002 steps.
"""
import numpy as np
import pytest
import torch

from laminarsmoke_amd import api
from laminarsmoke_amd.fields import workload_cells
from parity import (assert_parity_adaptive, assert_element_conservation, element_mass, oracle_envelope,
                    oracle_jacobian_envelope, reconcile_statuses, ORACLE_THREADS)

pytestmark = pytest.mark.gpu

RTOL, ATOL, DT = 1e-7, 1e-12, 1e-5
FAILURE_STATUSES = {-1, -2, -3, -4, -5, -6, -7, -8, -9}


def make_step(mech, jacobian):
    ode = api.ODE_Parameters()
    ode.SetRelativeTolerance(RTOL)
    ode.SetAbsoluteTolerance(ATOL)
    return api.ChemistryStep(mech, ode, method="team", jacobian=jacobian)


@pytest.fixture(scope="module")
def full_field(hip_device):
    mech, y = workload_cells(3)
    n = y.shape[1]
    assert n == 1 << 20
    print(f"cfg3: {n} cells, one call (team, analytic Jacobian)", flush=True)
    dev = torch.device("cuda:0")
    y_in = torch.from_numpy(y).to(dev)
    y_out = torch.empty_like(y_in)
    dts = torch.full((n,), DT, dtype=torch.float64, device=dev)
    st = torch.full((n,), -99, dtype=torch.int32, device=dev)
    q = torch.zeros(n, dtype=torch.float64, device=dev)
    cnt = torch.zeros((n, 12), dtype=torch.int32, device=dev)
    step = make_step(mech, "analytic")
    try:
        step.integrate_device(n, dts.data_ptr(), y_in.data_ptr(), y_out.data_ptr(), st.data_ptr(), q.data_ptr(),
                              cnt.data_ptr(), torch.cuda.current_stream().cuda_stream, pressure=101325.0)
        torch.cuda.synchronize()
    finally:
        step.close()
    out = dict(mech=mech, y=y, yg=y_out.cpu().numpy(), st=st.cpu().numpy(), q=q.cpu().numpy(), cnt=cnt.cpu().numpy())
    del y_in, y_out
    print(f"cfg3: statuses {dict(zip(*np.unique(out['st'], return_counts=True)))}; steps median "
          f"{np.median(out['cnt'][:, 0])} max {out['cnt'][:, 0].max()}", flush=True)
    return out


def test_full_field_properties(full_field):
    mech, y, yg, st, cnt = (full_field[k] for k in ("mech", "y", "yg", "st", "cnt"))
    ns = mech.ns
    assert set(np.unique(st)) <= {2} | FAILURE_STATUSES, np.unique(st)
    assert (st == 2).mean() > 0.999, dict(zip(*np.unique(st, return_counts=True)))
    assert np.isfinite(yg).all()
    assert (yg[-1] >= 200.0).all() and (yg[-1] <= 6000.0).all()
    assert (yg[:ns] >= 0.0).all()
    np.testing.assert_allclose(yg[:ns].sum(0), 1.0, rtol=0, atol=1e-14)
    ok = st == 2
    assert (cnt[ok, 0] >= 1).all() and (cnt[ok, 9] >= cnt[ok, 0]).all()     # steps, right-hand sides
    # failed cells return their clipped, renormalised input (T unchanged)
    fail = ~ok
    if fail.any():
        y0 = np.clip(y[:ns, fail], 0.0, None)
        np.testing.assert_allclose(yg[:ns, fail], y0 / y0.sum(0), rtol=1e-14, atol=1e-300)
        np.testing.assert_array_equal(yg[-1, fail], y[-1, fail])
    # element conservation on every cell, in chunks
    for a in range(0, y.shape[1], 1 << 17):
        sl = slice(a, a + (1 << 17))
        assert_element_conservation(mech, y[:, sl], yg[:, sl], cnt[sl, 10], what=f"cfg3 cells {a}+")
    dz = np.abs(element_mass(mech, yg[:, :4096]) - element_mass(mech, y[:, :4096] / y[:ns, :4096].sum(0))).max()
    print(f"cfg3 full field: {ok.sum()} integrated, {fail.sum()} failed; element drift (first 4096 cells) {dz:.2e}",
          flush=True)


def sample_cells(full_field):
    """64 cells stratified by temperature, the 8 costliest by steps, up to 8 failed cells."""
    y, st, cnt = full_field["y"], full_field["st"], full_field["cnt"]
    T = y[-1]
    rng = np.random.default_rng(3)
    edges = np.linspace(300.0, 2500.0, 9)
    idx = [rng.choice(np.where((T >= a) & (T < b))[0], 8, replace=False) for a, b in zip(edges[:-1], edges[1:])]
    idx.append(np.argsort(cnt[:, 0], kind="stable")[-8:])
    idx.append(np.where(st != 2)[0][:8])
    return np.unique(np.concatenate(idx))


@pytest.mark.parametrize("jacobian", ["analytic", "numerical"])
def test_sample_against_oracle(full_field, oracle, jacobian):
    mech = full_field["mech"]
    idx = sample_cells(full_field)
    y = np.ascontiguousarray(full_field["y"][:, idx])
    print(f"cfg3 parity ({jacobian} Jacobian) on {idx.size} cells", flush=True)
    yo, st_o, _, c_o = oracle.chemistry_step(mech, y, DT, rtol=RTOL, atol=ATOL, threads=ORACLE_THREADS)
    env = oracle_envelope(oracle, mech, y, DT, RTOL, ATOL, yo)
    if jacobian == "analytic":
        yg, st_g, cg = full_field["yg"][:, idx], full_field["st"][idx], full_field["cnt"][idx]
        env = np.maximum(env, oracle_jacobian_envelope(oracle, mech, y, DT, RTOL, ATOL, yo))
    else:
        step = make_step(mech, "numerical")
        try:
            yg, st_g, _, cg = step.integrate(y, DT)
        finally:
            step.close()
    if jacobian == "analytic":               # Newton matrix changed on purpose
        yo, st_o = reconcile_statuses(oracle, mech, y, DT, RTOL, ATOL, st_g, st_o, yo, what=f"cfg3/{jacobian}")
    assert np.array_equal(st_g, st_o), (st_g, st_o)
    (cell, frac), env = assert_parity_adaptive(oracle, mech, y, DT, RTOL, ATOL, yo, yg, env, what=f"cfg3/{jacobian}")
    assert_element_conservation(mech, y, yg, cg[:, 10], what=f"cfg3/{jacobian}", y_ref=yo)
    print(f"cfg3/{jacobian}: max ratio {cell.max():.3g} (envelope max {env.max():.3g}), components within 10x "
          f"{100 * frac:.2f}%; steps gpu/oracle median {np.median(cg[:, 0])}/{np.median(c_o[:, 0])}, max "
          f"{cg[:, 0].max()}/{c_o[:, 0].max()}", flush=True)
