"""GPU: batching machinery of the persistent integrators.

* a visiting order changes which wave takes which cell, never a cell's result:
  ordered and unordered calls agree bit for bit (team and wide kernels, numerical
  and analytic Jacobians; the cell loop is independent per cell, and the wide
  kernel's FP32 dR/dc scatter is ordered across its waves);
* lsm_cost_order returns a permutation sorted by the previous call's step counts;
* the autotuner keeps a valid launch shape and results stay bit-identical;
* a chemistry step captured in a HIP graph and replayed several times, with eager
  work in between, reproduces the eager result every time.
"""
import numpy as np
import pytest
import torch

from laminarsmoke_amd import api
from laminarsmoke_amd.fields import workload_cells

pytestmark = pytest.mark.gpu

RTOL, ATOL, DT = 1e-7, 1e-12, 1e-5


def make(mech, method, jacobian="numerical"):
    ode = api.ODE_Parameters()
    ode.SetRelativeTolerance(RTOL)
    ode.SetAbsoluteTolerance(ATOL)
    return api.ChemistryStep(mech, ode, method=method, jacobian=jacobian)


class Batch:
    def __init__(self, y):
        dev = torch.device("cuda:0")
        self.n = y.shape[1]
        self.y = torch.from_numpy(y).to(dev)
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
        return self.out.cpu().numpy().copy(), self.st.cpu().numpy().copy(), self.cnt.cpu().numpy().copy()


@pytest.mark.parametrize("method,cfg,n,jacobian", [("team", 2, 512, "numerical"), ("wide", 1, 256, "numerical"),
                                                   ("team", 2, 512, "analytic"), ("wide", 2, 256, "analytic")])
def test_order_does_not_change_results(hip_device, method, cfg, n, jacobian):
    mech, y = workload_cells(cfg, n=n)
    print(f"order test {method} {jacobian} cfg{cfg} n={n}", flush=True)
    step = make(mech, method, jacobian)
    b = Batch(y)
    s = torch.cuda.current_stream()
    try:
        b.run(step, s)
        y0, st0, c0 = b.result()
        order = torch.empty(n, dtype=torch.int32, device="cuda:0")
        step.cost_order(n, b.cnt.data_ptr(), order.data_ptr(), s.cuda_stream)
        torch.cuda.synchronize()
        o = order.cpu().numpy()
        assert np.array_equal(np.sort(o), np.arange(n))
        assert np.all(np.diff(c0[o, 0]) <= 0), "not sorted by steps, most first"
        b.run(step, s, order)
        y1, st1, c1 = b.result()
        rev = torch.flip(order, [0]).contiguous()
        b.run(step, s, rev)
        y2, st2, c2 = b.result()
    finally:
        step.close()
    for ya, sa, ca in ((y1, st1, c1), (y2, st2, c2)):
        assert np.array_equal(ya, y0) and np.array_equal(sa, st0) and np.array_equal(ca, c0)
    print(f"{method}: ordered results identical; steps range {c0[:, 0].min()}-{c0[:, 0].max()}", flush=True)


def test_autotune_keeps_results(hip_device):
    mech, y = workload_cells(2, n=2048)
    print("autotune test cfg2 n=2048", flush=True)
    step = make(mech, "team")
    b = Batch(y)
    s = torch.cuda.current_stream()
    try:
        b.run(step, s)
        y0, st0, c0 = b.result()
        before = step.launch()
        res = step.autotune(b.n, b.dt.data_ptr(), b.y.data_ptr(), 1024, s.cuda_stream)
        after = step.launch()
        b.run(step, s)
        y1, st1, c1 = b.result()
    finally:
        step.close()
    assert 1 <= res["waves_per_block"] <= 4 and res["grid"] > 0 and res["best_us"] > 0
    assert after["waves_per_block"] == res["waves_per_block"] and after["grid"] == res["grid"]
    assert np.array_equal(y1, y0) and np.array_equal(st1, st0)
    print(f"autotune: {before} -> {after} ({res['best_us']} us on the sample)", flush=True)


def test_graph_replay_matches_eager(hip_device):
    mech, y = workload_cells(1, n=1024)
    print("graph test cfg1 n=1024", flush=True)
    step = make(mech, "team")
    b = Batch(y)
    s = torch.cuda.current_stream()
    try:
        b.run(step, s)
        y0, st0, c0 = b.result()
        print("graph: eager step done", flush=True)
        side = torch.cuda.Stream()
        side.wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(side):
            with torch.cuda.graph(g, stream=side):
                b.run(step, side)
        s.wait_stream(side)
        torch.cuda.synchronize()
        print("graph: captured", flush=True)
        for rep in range(3):
            b.out.zero_()
            b.st.fill_(-99)
            g.replay()
            y1, st1, c1 = b.result()
            print(f"graph: replay {rep} done", flush=True)
            assert np.array_equal(y1, y0) and np.array_equal(st1, st0) and np.array_equal(c1, c0), rep
            b.run(step, s)                      # eager work between replays
            y2, _, _ = b.result()
            assert np.array_equal(y2, y0)
    finally:
        step.close()
    print("graph: 3 replays identical to eager", flush=True)


@pytest.mark.parametrize("method", ["team", "wide"])
@pytest.mark.parametrize("jac", ["numerical", "analytic"])
def test_repeated_launches_identical(hip_device, method, jac):
    """Every cell starts from a fresh solver: the same cells give bit-identical results on every
    launch, whatever the wave or workgroup that took them last time (a solver array read before
    it was written once made the first launch differ on 5 % of the 14-species cells)."""
    mech, y = workload_cells(1, n=1024)
    print(f"repeat test {method} {jac} cfg1 n=1024", flush=True)
    step = make(mech, method, jac)
    b = Batch(y)
    s = torch.cuda.current_stream()
    try:
        b.run(step, s)
        y0, st0, c0 = b.result()
        for rep in range(2):
            b.out.zero_()
            b.run(step, s)
            y1, st1, c1 = b.result()
            assert np.array_equal(y1, y0) and np.array_equal(st1, st0) and np.array_equal(c1, c0), rep
    finally:
        step.close()
