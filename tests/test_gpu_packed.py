"""GPU: the team kernel with several cells per wavefront (lsm_set_cells_per_wave).

With 2 (ns+1 <= 32) or 4 (ns+1 <= 16) cells per 64-lane wavefront each cell runs on its
own 32- or 16-lane segment.  Its arithmetic is the one-cell-per-wavefront kernel's; only
the order of the formation-rate scatter changes (reactions are strided over the segment's
lanes), so results are compared with the oracle, under the criteria of tests/parity.py
(every cell of the 4,096-cell 14-species field, 4 and 2 cells per wavefront, is in
test_gpu_parity.py::test_cfg1_whole_field).  Here:

* a cell's result does not depend on the segment or wavefront it lands in: bit-identical
  under any visiting order, for batches that leave segments of the last wavefront empty,
  and for one cell alone;
* HIP-graph replay with eager work in between reproduces the eager result;
* the autotuner's search over cells per wavefront keeps a valid, reproducible choice;
* packings the mechanism does not allow are refused.

Per-case parity of the packed kernel (14-species field, t0 != 0, the temperature gate,
statuses -1 / -2 / -9, constant volume, compact output, the reference's CHEMKIN mechanisms)
is in test_gpu_parity.py, test_gpu_cell_loop.py, test_gpu_reactor_modes.py and
test_gpu_chemkin.py as the methods "pack4" / "pack2".
"""
import os

import numpy as np
import pytest
import torch

from laminarsmoke_amd import api
from laminarsmoke_amd.chemkin import load_chemkin
from laminarsmoke_amd.fields import workload_cells, premixed_cells
from laminarsmoke_amd.mechanism import workload_mechanism
from parity import assert_parity_adaptive, assert_element_conservation, oracle_envelope, ORACLE_THREADS

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RTOL, ATOL, DT = 1e-7, 1e-12, 1e-5


def make(mech, cpw, jacobian="numerical"):
    ode = api.ODE_Parameters()
    ode.SetRelativeTolerance(RTOL)
    ode.SetAbsoluteTolerance(ATOL)
    return api.ChemistryStep(mech, ode, method="team", jacobian=jacobian, cells_per_wave=cpw)


class Batch:
    def __init__(self, y):
        dev = torch.device("cuda:0")
        self.n = y.shape[1]
        self.y = torch.from_numpy(np.ascontiguousarray(y)).to(dev)
        self.dt = torch.full((self.n,), DT, dtype=torch.float64, device=dev)
        self.out = torch.empty_like(self.y)
        self.st = torch.zeros(self.n, dtype=torch.int32, device=dev)
        self.q = torch.zeros(self.n, dtype=torch.float64, device=dev)
        self.cnt = torch.zeros((self.n, 12), dtype=torch.int32, device=dev)

    def run(self, step, stream, order=None):
        step.integrate_device(self.n, self.dt.data_ptr(), self.y.data_ptr(), self.out.data_ptr(), self.st.data_ptr(),
                              self.q.data_ptr(), self.cnt.data_ptr(), stream.cuda_stream,
                              order.data_ptr() if order is not None else None)

    def result(self):
        torch.cuda.synchronize()
        return (self.out.cpu().numpy().copy(), self.st.cpu().numpy().copy(), self.cnt.cpu().numpy().copy(),
                self.q.cpu().numpy().copy())


@pytest.mark.parametrize("cpw,jac", [(4, "numerical"), (4, "analytic"), (2, "analytic")])
def test_segment_position_does_not_change_results(hip_device, cpw, jac):
    """61 cells (the last wavefront has empty segments): natural, reversed and shuffled visiting
    orders, and single cells alone, give bit-identical results."""
    mech, y = workload_cells(1, n=61)
    n = y.shape[1]
    print(f"packed order test cpw={cpw} jac={jac} n={n}", flush=True)
    step = make(mech, cpw, jac)
    s = torch.cuda.current_stream()
    b = Batch(y)
    try:
        b.run(step, s)
        ref = b.result()
        rev = torch.arange(n - 1, -1, -1, dtype=torch.int32, device="cuda:0")
        perm = torch.from_numpy(np.random.default_rng(3).permutation(n).astype(np.int32)).to("cuda:0")
        for order in (rev, perm):
            b.out.zero_()
            b.st.fill_(-99)
            b.run(step, s, order)
            got = b.result()
            for a, r in zip(got, ref):
                assert np.array_equal(a, r)
        for c in (0, 29, 60):     # one cell alone: segment 0 of the only wavefront
            one = Batch(y[:, c:c + 1])
            one.run(step, s)
            got = one.result()
            assert np.array_equal(got[0][:, 0], ref[0][:, c]) and got[1][0] == ref[1][c]
            assert np.array_equal(got[2][0], ref[2][c]) and got[3][0] == ref[3][c]
    finally:
        step.close()
    print(f"packed @{cpw}/{jac}: identical in every order and alone; steps {ref[2][:, 0].min()}-"
          f"{ref[2][:, 0].max()}", flush=True)


def test_graph_replay_matches_eager(hip_device):
    mech, y = workload_cells(1, n=1024)
    print("packed graph test cfg1 n=1024 cpw=4", flush=True)
    step = make(mech, 4)
    b = Batch(y)
    s = torch.cuda.current_stream()
    try:
        b.run(step, s)
        ref = b.result()
        side = torch.cuda.Stream()
        side.wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(side):
            with torch.cuda.graph(g, stream=side):
                b.run(step, side)
        s.wait_stream(side)
        torch.cuda.synchronize()
        for rep in range(3):
            b.out.zero_()
            b.st.fill_(-99)
            g.replay()
            got = b.result()
            assert all(np.array_equal(a, r) for a, r in zip(got, ref)), rep
            b.run(step, s)                      # eager work between replays
            assert np.array_equal(b.result()[0], ref[0])
    finally:
        step.close()
    print("packed graph: 3 replays identical to eager", flush=True)


def test_autotune_searches_cells_per_wave(hip_device):
    mech, y = workload_cells(1)
    print("packed autotune test cfg1 n=4096", flush=True)
    step = make(mech, 1, "analytic")
    b = Batch(y)
    s = torch.cuda.current_stream()
    try:
        res = step.autotune(b.n, b.dt.data_ptr(), b.y.data_ptr(), 4096, s.cuda_stream)
        launch = step.launch()
        b.run(step, s)
        got = b.result()
    finally:
        step.close()
    assert res["cells_per_wave"] in (1, 2, 4) and launch["cells_per_wave"] == res["cells_per_wave"]
    assert 1 <= res["waves_per_block"] <= 4 and res["grid"] > 0 and res["best_us"] > 0
    assert launch["waves_per_block"] == res["waves_per_block"] and launch["grid"] == res["grid"]
    # the tuned handle gives what a fresh handle with the same packing gives
    fresh = make(mech, res["cells_per_wave"], "analytic")
    b2 = Batch(y)
    try:
        b2.run(fresh, s)
        want = b2.result()
    finally:
        fresh.close()
    assert all(np.array_equal(a, r) for a, r in zip(got, want))
    print(f"packed autotune: {res}", flush=True)


def test_packings_the_mechanism_does_not_allow(hip_device):
    ch4 = load_chemkin(os.path.join(GOLDEN, "POLIMI_CH4_SKELETAL_1412.CKI"),
                       os.path.join(GOLDEN, "POLIMI_CH4_SKELETAL_1412_thermo.CKT"))   # ns+1 = 25
    step = make(ch4, 2)
    try:
        assert step.cells_per_wave() == 2
        with pytest.raises(api.LsmError):
            step.set_cells_per_wave(4)
        for bad in (0, 3, 8, -1):
            with pytest.raises(api.LsmError):
                step.set_cells_per_wave(bad)
        assert step.cells_per_wave() == 2
        step.set_cells_per_wave(1)
        assert step.cells_per_wave() == 1
    finally:
        step.close()
    with pytest.raises(api.LsmError):
        make(workload_mechanism(53), 2)          # ns+1 = 54


def test_ch4_skeletal_two_cells_per_wave(hip_device, oracle):
    """The reference's CH4 skeletal mechanism (24 species, SRI fall-off) at the tutorial's
    tolerance, two cells per wavefront, 256 premixed cells against the oracle."""
    mech = load_chemkin(os.path.join(GOLDEN, "POLIMI_CH4_SKELETAL_1412.CKI"),
                        os.path.join(GOLDEN, "POLIMI_CH4_SKELETAL_1412_thermo.CKT"))
    y = premixed_cells(mech, 256, seed=11)
    dt, rtol, atol = 1e-4, 1e-6, 1e-12
    print("packed CH4 skeletal cpw=2 cells=256", flush=True)
    yo, st_o, q_o, c_o = oracle.chemistry_step(mech, y, dt, rtol=rtol, atol=atol, threads=ORACLE_THREADS)
    ode = api.ODE_Parameters()
    ode.SetRelativeTolerance(rtol)
    ode.SetAbsoluteTolerance(atol)
    step = api.ChemistryStep(mech, ode, method="team", cells_per_wave=2)
    try:
        yg, st_g, q_g, c_g = step.integrate(y, dt)
    finally:
        step.close()
    assert np.array_equal(st_g, st_o), (st_g, st_o)
    env = oracle_envelope(oracle, mech, y, dt, rtol, atol, yo)
    (cell, frac), env = assert_parity_adaptive(oracle, mech, y, dt, rtol, atol, yo, yg, env, what="ch4/@2")
    assert_element_conservation(mech, y, yg, c_g[:, 10], what="ch4/@2", y_ref=yo)
    ignited = int((yo[mech.ns] - y[mech.ns] > 100.0).sum())
    print(f"ch4/@2: max ratio {cell.max():.3g} (envelope max {env.max():.3g}); {ignited} of 256 ignite; steps "
          f"gpu/oracle median {np.median(c_g[:, 0])}/{np.median(c_o[:, 0])}", flush=True)
