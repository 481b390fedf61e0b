"""GPU: serving chemistry steps from several CFD cases (include/lsm_serve.h).

Three cases with different mechanisms (14, 53 and 160 species), mesh sizes and
time steps each run two chemistry steps alone, then all together on one server
whose HBM budget cannot hold every case at once (cases are evicted to host memory
and brought back), with their steps overlapping on the device.  Each case's
final state must be the same alone and in the mix: bit for bit (the FP64 team and
the wide kernel with the numerical Jacobian integrate every cell independently of
what else runs), which is stricter than the stated tolerance.  The states also
equal two plain ChemistryStep calls.
"""
import numpy as np
import pytest

from laminarsmoke_amd import api
from laminarsmoke_amd.fields import workload_cells

pytestmark = pytest.mark.gpu

CASES = [(1, 2000, (1e-5, 3e-6)), (2, 1000, (1e-5, 2e-5)), (4, 48, (1e-6, 1e-6))]


def make_step(cfg):
    ode = api.ODE_Parameters()
    ode.SetRelativeTolerance(1e-6 if cfg == 4 else 1e-7)
    ode.SetAbsoluteTolerance(1e-12)
    return api.ChemistryStep(workload_cells(cfg, n=8)[0], ode, method="auto")


def case_inputs():
    return [(cfg, workload_cells(cfg, n=n)[1], dts) for cfg, n, dts in CASES]


def run_server(inputs, budget, max_inflight):
    srv = api.Server(budget, max_inflight)
    steps, ids = [], []
    try:
        for cfg, y, dts in inputs:
            st = make_step(cfg)
            steps.append(st)
            ids.append(srv.add_case(st, y))
        for k in range(2):                               # interleaved submissions
            for cid, (cfg, y, dts) in zip(ids, inputs):
                srv.submit(cid, dts[k])
        srv.drain()
        done = srv.poll()
        out = [srv.state(cid) for cid in ids]
        stats = srv.stats()
    finally:
        srv.close()
        for st in steps:
            st.close()
    return out, done, stats


def test_same_alone_and_in_the_mix(hip_device):
    inputs = case_inputs()
    print("serving: alone", flush=True)
    alone = []
    for inp in inputs:
        out, done, stats = run_server([inp], 8 << 30, 1)
        assert len(done) == 2 and stats["evictions"] == 0
        alone.append(out[0])
    # direct calls give the same states
    for (cfg, y, dts), ya in zip(inputs, alone):
        st = make_step(cfg)
        try:
            y1, s1, _, _ = st.integrate(y, dts[0])
            y2, s2, _, _ = st.integrate(y1, dts[1])
        finally:
            st.close()
        assert np.array_equal(y2, ya), f"cfg{cfg}: server differs from direct calls"
    # the mix, with a budget that holds the two largest footprints but not all three
    sizes = [y.shape[1] * (8 * (y.shape[0] + 1) + 4 * 14) for cfg, y, dts in inputs]   # state bytes per case
    budget = int(sum(sorted(sizes)[-2:]) * 1.05)
    print(f"serving: mix, footprints {sizes}, budget {budget}", flush=True)
    mixed, done, stats = run_server(inputs, budget, 3)
    assert len(done) == 6
    assert stats["evictions"] > 0, stats
    assert stats["peak_bytes"] <= budget, stats
    for (cfg, y, dts), ya, ym in zip(inputs, alone, mixed):
        assert np.array_equal(ym, ya), f"cfg{cfg}: result in the mix differs from alone"
    lat = sorted(d[1] for d in done)
    print(f"serving: identical alone and in the mix; stats {stats}; latencies {['%.3f' % v for v in lat]} s",
          flush=True)
