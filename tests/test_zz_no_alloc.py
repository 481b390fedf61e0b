"""GPU: lsm_integrate_batch_device_noalloc, the chemistry step that allocates nothing, through the
C-ABI against lsm_integrate_batch_device_ordered (the entry point it varies) on the same cells.

The two run the same kernels; only where their scratch lives differs (the library's static
per-GPU array against the handle's buffers), and the wide and plain kernels may run on fewer
workgroups / cell slots.  Every cell is integrated independently of the others, so states,
statuses, Q and counters must be bit-identical.  Two calls in a row, the second on fewer cells
and in another visiting order (the array keeps the first call's contents, so whatever must start
at zero has to be cleared by the call), for the team, wide and plain kernels on the 14-, 53- and
160-species mechanisms.  The second call must leave the device's free memory unchanged.
"""
import numpy as np
import pytest
import torch

from laminarsmoke_amd import api
from laminarsmoke_amd.fields import workload_cells

pytestmark = pytest.mark.gpu

DT, P = 1e-5, 101325.0

# (kernel, workload, cells of the first call, cells of the second call)
CASES = [("team", 1, 256, 100), ("team", 2, 96, 40), ("wide", 2, 24, 9), ("plain", 1, 16, 7), ("wide", 4, 8, 5)]


def make_step(mech, method, cfg):
    ode = api.ODE_Parameters()
    ode.SetRelativeTolerance(1e-6 if cfg == 4 else 1e-7)
    ode.SetAbsoluteTolerance(1e-12)
    return api.ChemistryStep(mech, ode, method=method)


class Buffers:
    """Device inputs and outputs of one call (allocated before any timed or measured call)."""

    def __init__(self, y, order):
        dev = torch.device("cuda:0")
        n = y.shape[1]
        self.n = n
        self.y_in = torch.from_numpy(np.ascontiguousarray(y)).to(dev)
        self.y_out = torch.full_like(self.y_in, float("nan"))
        self.dt = torch.full((n,), DT, dtype=torch.float64, device=dev)
        self.st = torch.full((n,), -99, dtype=torch.int32, device=dev)
        self.q = torch.full((n,), -1.0, dtype=torch.float64, device=dev)
        self.cnt = torch.full((n, 12), -1, dtype=torch.int32, device=dev)
        self.order = None if order is None else torch.from_numpy(order.astype(np.int32)).to(dev)

    def args(self):
        return (self.n, 0.0, P, self.dt.data_ptr(), self.y_in.data_ptr(), self.y_out.data_ptr(), self.st.data_ptr(),
                self.q.data_ptr(), self.cnt.data_ptr(), None if self.order is None else self.order.data_ptr())

    def result(self):
        return (self.y_out.cpu().numpy(), self.st.cpu().numpy(), self.q.cpu().numpy(), self.cnt.cpu().numpy())


def call(step, b, noalloc):
    n, t0, p, dt, yi, yo, st, q, cnt, order = b.args()
    stream = torch.cuda.current_stream().cuda_stream
    if noalloc:
        rc = api.lib().lsm_integrate_batch_device_noalloc(step.h, n, t0, p, dt, yi, yo, st, q, cnt, order, stream)
    elif order is not None:
        rc = api.lib().lsm_integrate_batch_device_ordered(step.h, n, t0, p, dt, yi, yo, st, q, cnt, order, stream)
    else:
        rc = api.lib().lsm_integrate_batch_device(step.h, n, t0, p, dt, yi, yo, st, q, cnt, stream)
    assert rc == 0, api.lib().lsm_last_error().decode()


def assert_same(a, b, what):
    for name, x, y in zip(("y", "status", "q", "counters"), a, b):
        assert np.array_equal(x, y, equal_nan=False), f"{what}: {name} differs"


@pytest.mark.parametrize("method,cfg,n1,n2", CASES)
def test_noalloc_matches_ordered_entry(hip_device, method, cfg, n1, n2):
    mech, y = workload_cells(cfg, n=n1)
    y2 = np.ascontiguousarray(y[:, ::-1][:, :n2])          # other cells of the set, fewer
    order2 = np.arange(n2)[::-1].copy()                      # and visited in another order
    print(f"noalloc {method} cfg{cfg}: calls of {n1} then {n2} cells", flush=True)
    step = make_step(mech, method, cfg)
    try:
        ref1, ref2 = Buffers(y, None), Buffers(y2, order2)
        na1, na2 = Buffers(y, None), Buffers(y2, order2)
        call(step, ref1, False)
        call(step, ref2, False)
        torch.cuda.synchronize()
        call(step, na1, True)                                # first call on this device: set-up
        torch.cuda.synchronize()
        free0, _ = torch.cuda.mem_get_info()
        call(step, na2, True)
        torch.cuda.synchronize()
        free1, _ = torch.cuda.mem_get_info()
        assert step.launch()["method"] == method
    finally:
        step.close()
    r1, r2 = ref1.result(), ref2.result()
    assert (r1[1] != -99).all() and (r2[1] != -99).all()
    assert_same(na1.result(), r1, f"{method}/cfg{cfg} first call")
    assert_same(na2.result(), r2, f"{method}/cfg{cfg} second call")
    assert free1 == free0, (free0, free1)
    print(f"noalloc {method} cfg{cfg}: identical on both calls; statuses {np.unique(r1[1])}", flush=True)


def test_noalloc_rejections(hip_device):
    mech, y = workload_cells(1, n=4)
    b = Buffers(y, None)
    step = make_step(mech, "team", 1)
    try:
        n, t0, p, dt, yi, yo, st, q, cnt, order = b.args()
        stream = torch.cuda.current_stream().cuda_stream
        L = api.lib()
        assert L.lsm_integrate_batch_device_noalloc(step.h, -1, t0, p, dt, yi, yo, st, q, cnt, None, stream) != 0
        assert L.lsm_integrate_batch_device_noalloc(step.h, n, float("nan"), p, dt, yi, yo, st, q, cnt, None,
                                                    stream) != 0
        assert L.lsm_integrate_batch_device_noalloc(step.h, n, t0, p, dt, None, yo, st, q, cnt, None, stream) != 0
        assert L.lsm_integrate_batch_device_noalloc(step.h, 0, t0, p, dt, yi, yo, st, q, cnt, None, stream) == 0
    finally:
        step.close()
    cv = api.ChemistryStep(mech, api.ODE_Parameters(), method="team", constant_volume=True)
    try:
        rc = api.lib().lsm_integrate_batch_device_noalloc(cv.h, n, t0, p, dt, yi, yo, st, q, cnt, None,
                                                          torch.cuda.current_stream().cuda_stream)
        assert rc != 0 and b"constant-volume" in api.lib().lsm_last_error()
    finally:
        cv.close()
