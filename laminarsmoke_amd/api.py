"""Python view of the C-ABI (include/laminarsmoke_amd.h), for tests and benchmarks.

The names mirror the reference interface of the chemistry step:

* :class:`ODE_Parameters` - OpenSMOKE::ODE_Parameters
  (math/external-ode-solvers/ODE_Parameters.h:52-160): SetType,
  SetRelativeTolerance, SetAbsoluteTolerance, SetFullPivoting, SetMaximumOrder,
  SetMaximumNumberOfSteps, SetupFromDictionary;
* :class:`ChemistryStep` - the OpenSMOKE branch of the cell loop
  (LS/unsteady/chemistry_DI.H:76-331) on the GPU: one call integrates every cell.

Everything runs in liblsm_hip.so; this module only moves pointers.  Loading fails
loudly when the library is missing.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._abi import LsmOdeParameters, LsmReactorOptions, N_COUNTERS, COUNTER_NAMES

LIB_PATH = os.environ.get("LSM_LIB_PATH") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "liblsm_hip.so")

ODE_INTEGRATOR = {"OpenSMOKE": 0, "BzzOde": 1, "CVODE": 2, "DVODE": 3, "DASPK": 4, "DLSODE": 5, "DLSODA": 6,
                  "RADAU5": 7, "MEBDF": 8, "CHEMEQ2": 9, "OpenSMOKE_GPU": 100}
METHODS = {"auto": 0, "plain": 1, "team": 2, "wide": 3}
JACOBIANS = {"numerical": 0, "analytic": 1}
STATUS_NAMES = {0: "skipped", 2: "ok", -1: "max steps", -2: "too strict tolerances", -4: "error-test failures",
                -5: "convergence failures", -6: "step too small", -9: "illegal constraints"}

_lib = None


class LsmError(RuntimeError):
    pass


def lib():
    """The loaded HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LsmError(f"{LIB_PATH} is missing: build it with `python -m laminarsmoke_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64 = ctypes.c_void_p, ctypes.c_int64
        L.lsm_last_error.restype = ctypes.c_char_p
        L.lsm_version.restype = ctypes.c_char_p
        L.lsm_ode_parameters_default.argtypes = [vp]
        L.lsm_ode_parameters_parse.argtypes = [ctypes.c_char_p, vp]
        L.lsm_reactor_options_default.argtypes = [vp]
        L.lsm_create.argtypes = [vp, vp, vp, ctypes.POINTER(vp)]
        L.lsm_set_method.argtypes = [vp, ctypes.c_int32]
        L.lsm_set_cells_per_wave.argtypes = [vp, ctypes.c_int32]
        L.lsm_get_cells_per_wave.argtypes = [vp, ctypes.POINTER(ctypes.c_int32)]
        L.lsm_set_jacobian.argtypes = [vp, ctypes.c_int32]
        L.lsm_set_refinement.argtypes = [vp, ctypes.c_int32]
        L.lsm_fp64_fallbacks.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
        L.lsm_equations.argtypes = [vp, i64, vp, vp, vp]
        L.lsm_equations_device.argtypes = [vp, i64, vp, vp, vp, vp]
        L.lsm_destroy.argtypes = [vp]
        f64 = ctypes.c_double
        L.lsm_integrate_batch.argtypes = [vp, i64, f64, f64, vp, vp, vp, vp, vp, vp]
        L.lsm_integrate_batch_device.argtypes = [vp, i64, f64, f64, vp, vp, vp, vp, vp, vp, vp]
        L.lsm_integrate_batch_device_ordered.argtypes = [vp, i64, f64, f64, vp, vp, vp, vp, vp, vp, vp, vp]
        L.lsm_integrate_batch_rho.argtypes = [vp, i64, f64, f64, vp, vp, vp, vp, vp, vp, vp]
        L.lsm_properties.argtypes = [vp, i64, vp, vp, vp, vp, vp]
        L.lsm_properties_device.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp]
        L.lsm_set_transport.argtypes = [vp, vp]
        L.lsm_set_transport_options.argtypes = [vp, vp]
        L.lsm_transport_properties.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp]
        L.lsm_transport_properties_device.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, vp]
        L.lsm_integrate_batch_device_rho.argtypes = [vp, i64, f64, f64, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.lsm_integrate_batch_device_noalloc.argtypes = [vp, i64, f64, f64, vp, vp, vp, vp, vp, vp, vp, vp]
        L.lsm_cost_order.argtypes = [vp, i64, vp, vp, vp]
        L.lsm_device_bytes.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
        L.lsm_autotune.argtypes = [vp, i64, f64, f64, vp, vp, i64, vp, vp]
        L.lsm_get_launch.argtypes = [vp, vp]
        L.lsm_server_create.argtypes = [ctypes.c_size_t, ctypes.c_int32, ctypes.POINTER(vp)]
        L.lsm_server_destroy.argtypes = [vp]
        L.lsm_server_add_case.argtypes = [vp, vp, i64, vp, ctypes.POINTER(ctypes.c_int32)]
        L.lsm_server_submit.argtypes = [vp, ctypes.c_int32, f64, f64, vp, ctypes.POINTER(ctypes.c_uint64)]
        L.lsm_server_poll.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
        L.lsm_server_drain.argtypes = [vp]
        L.lsm_server_read_state.argtypes = [vp, ctypes.c_int32, vp]
        L.lsm_server_stats.argtypes = [vp, vp]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise LsmError(f"lsm error {rc}: {lib().lsm_last_error().decode()}")


class ODE_Parameters:
    """OpenSMOKE::ODE_Parameters (ODE_Parameters.h:52-160) on top of lsm_ode_parameters."""

    def __init__(self):
        self.c = LsmOdeParameters()
        _check(lib().lsm_ode_parameters_default(ctypes.byref(self.c)))

    def SetType(self, name_or_value):
        self.c.type = ODE_INTEGRATOR[name_or_value] if isinstance(name_or_value, str) else int(name_or_value)

    def SetRelativeTolerance(self, v):
        self.c.relative_tolerance = float(v)

    def SetAbsoluteTolerance(self, v):
        self.c.absolute_tolerance = float(v)

    def SetFullPivoting(self, flag):
        self.c.full_pivoting = int(bool(flag))

    def SetMaximumOrder(self, v):
        self.c.maximum_order = int(v)

    def SetMaximumNumberOfSteps(self, v):
        self.c.maximum_number_of_steps = int(v)

    def SetupFromDictionary(self, text: str):
        _check(lib().lsm_ode_parameters_parse(text.encode(), ctypes.byref(self.c)))

    def type(self):
        return self.c.type

    def relative_tolerance(self):
        return self.c.relative_tolerance

    def absolute_tolerance(self):
        return self.c.absolute_tolerance

    def full_pivoting(self):
        return bool(self.c.full_pivoting)

    def maximum_order(self):
        return self.c.maximum_order


def reactor_options(pressure=101325.0, min_temperature=0.0, energy=True, mass_fractions_tol=1e-4,
                    constant_volume=False, compact=False):
    """lsm_reactor_options; pressure: the handle's default pressure (chemistry steps take theirs per
    call); constant_volume: constant-volume reactor; compact: Strang "compact" output (y_out holds the
    sources RR and RT instead of the final state)."""
    r = LsmReactorOptions()
    _check(lib().lsm_reactor_options_default(ctypes.byref(r)))
    r.pressure, r.min_temperature, r.energy_equation = pressure, min_temperature, int(bool(energy))
    r.mass_fractions_tol = mass_fractions_tol
    r.constant_volume, r.strang_compact = int(bool(constant_volume)), int(bool(compact))
    return r


class ChemistryStep:
    """Batched chemistry step of one mechanism on the current HIP device."""

    def __init__(self, mech, ode: ODE_Parameters | None = None, method="auto", jacobian="numerical",
                 cells_per_wave=1, **reactor_kw):
        self.mech = mech
        self.ode = ode or ODE_Parameters()
        self.reactor = reactor_options(**reactor_kw)
        h = ctypes.c_void_p()
        _check(lib().lsm_create(ctypes.byref(mech.struct()), ctypes.byref(self.ode.c),
                                ctypes.byref(self.reactor), ctypes.byref(h)))
        self.h = h
        try:
            self.set_method(method)
            self.set_jacobian(jacobian)
            if cells_per_wave != 1:
                self.set_cells_per_wave(cells_per_wave)
        except LsmError:
            self.close()
            raise

    def set_method(self, method):
        _check(lib().lsm_set_method(self.h, METHODS[method] if isinstance(method, str) else int(method)))

    def set_cells_per_wave(self, cells):
        """Cells per wavefront of the team kernel: 1, 2 (ns+1 <= 32) or 4 (ns+1 <= 16)."""
        _check(lib().lsm_set_cells_per_wave(self.h, int(cells)))

    def cells_per_wave(self):
        v = ctypes.c_int32()
        _check(lib().lsm_get_cells_per_wave(self.h, ctypes.byref(v)))
        return v.value

    def set_refinement(self, sweeps):
        _check(lib().lsm_set_refinement(self.h, int(sweeps)))

    def fp64_fallbacks(self):
        """Wide-kernel factorisations redone in FP64 since creation."""
        v = ctypes.c_uint64()
        _check(lib().lsm_fp64_fallbacks(self.h, ctypes.byref(v)))
        return v.value

    def set_jacobian(self, kind):
        _check(lib().lsm_set_jacobian(self.h, JACOBIANS[kind] if isinstance(kind, str) else int(kind)))

    def equations(self, y, jacobian=True):
        """Reactor RHS f[ns+1, n] and (optionally) analytic Jacobian J[n, ns+1, ns+1] at states y[ns+1, n]."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        neq, n = y.shape
        if neq != self.mech.neq:
            raise ValueError(f"y has {neq} rows, mechanism needs {self.mech.neq}")
        f = np.empty_like(y)
        J = np.empty((n, neq, neq)) if jacobian else None
        _check(lib().lsm_equations(self.h, n, y.ctypes.data, f.ctypes.data, J.ctypes.data if J is not None else None))
        return f, J

    def close(self):
        if self.h:
            lib().lsm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _p(self, pressure):
        return float(self.reactor.pressure if pressure is None else pressure)

    def integrate(self, y, dt, q=None, with_counters=True, rho=None, t0=0.0, pressure=None):
        """Host arrays: y[ns+1, n] -> (y_out, status, q, counters[n, 12]).  t0: start time of the step
        (t - dt of the CFD step); pressure: reactor pressure [Pa] (default: the handle's); rho: per-cell
        densities of a constant-volume reactor (default: ideal-gas density of the input state)."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        neq, n = y.shape
        if neq != self.mech.neq:
            raise ValueError(f"y has {neq} rows, mechanism needs {self.mech.neq}")
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (n,)))
        yout = np.empty_like(y)
        st = np.zeros(n, dtype=np.int32)
        qa = np.zeros(n) if q is None else np.ascontiguousarray(q, dtype=np.float64).copy()
        cnt = np.zeros((n, N_COUNTERS), dtype=np.int32) if with_counters else None
        p = lambda a: a.ctypes.data if a is not None else None
        if rho is not None:
            rho = np.ascontiguousarray(np.broadcast_to(np.asarray(rho, dtype=np.float64), (n,)))
        _check(lib().lsm_integrate_batch_rho(self.h, n, float(t0), self._p(pressure), p(dt), p(rho), p(y), p(yout),
                                             p(st), p(qa), p(cnt)))
        return yout, st, qa, cnt

    PROPS = ("MW", "cTot", "psi", "h", "cp", "cv")

    def properties(self, y, p=None, species=True):
        """Host arrays: y[ns+1, n] -> (props[6, n] rows MW, cTot, psi, h, cp, cv; x[ns, n]; cp_k[ns, n])."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        neq, n = y.shape
        if neq != self.mech.neq:
            raise ValueError(f"y has {neq} rows, mechanism needs {self.mech.neq}")
        out = np.empty((6, n))
        x = np.empty((self.mech.ns, n)) if species else None
        cpk = np.empty((self.mech.ns, n)) if species else None
        if p is not None:
            p = np.ascontiguousarray(np.broadcast_to(np.asarray(p, dtype=np.float64), (n,)))
        q = lambda a: a.ctypes.data if a is not None else None
        _check(lib().lsm_properties(self.h, n, q(y), q(p), q(out), q(x), q(cpk)))
        return out, x, cpk

    def properties_device(self, n, y_ptr, props_ptr, p_ptr=None, x_ptr=None, cpk_ptr=None, stream=None):
        _check(lib().lsm_properties_device(self.h, int(n), y_ptr, p_ptr, props_ptr, x_ptr, cpk_ptr, stream))

    def set_transport(self, tr):
        """Copy the transport fits (laminarsmoke_amd.transport.Transport) to the device."""
        self._transport = tr
        _check(lib().lsm_set_transport(self.h, ctypes.byref(tr.struct())))

    def set_transport_options(self, opts=None):
        """The physical-model options of the refresh (laminarsmoke_amd.transport.TransportOptions);
        None restores the defaults."""
        self._transport_opts = opts
        _check(lib().lsm_set_transport_options(self.h, ctypes.byref(opts.struct()) if opts is not None else None))

    def transport(self, y, p=None, species=True):
        """Host arrays: y[ns+1, n] -> (mu[n], lambda[n], dmix[ns, n], teta[ns, n]) (the last two None
        unless `species`)."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        neq, n = y.shape
        if neq != self.mech.neq:
            raise ValueError(f"y has {neq} rows, mechanism needs {self.mech.neq}")
        mu, lam = np.empty(n), np.empty(n)
        dmix = np.empty((self.mech.ns, n)) if species else None
        teta = np.empty((self.mech.ns, n)) if species else None
        if p is not None:
            p = np.ascontiguousarray(np.broadcast_to(np.asarray(p, dtype=np.float64), (n,)))
        q = lambda a: a.ctypes.data if a is not None else None
        _check(lib().lsm_transport_properties(self.h, n, q(y), q(p), q(mu), q(lam), q(dmix), q(teta)))
        return mu, lam, dmix, teta

    def transport_device(self, n, y_ptr, mu_ptr, lambda_ptr, p_ptr=None, dmix_ptr=None, teta_ptr=None, stream=None):
        _check(lib().lsm_transport_properties_device(self.h, int(n), y_ptr, p_ptr, mu_ptr, lambda_ptr, dmix_ptr,
                                                     teta_ptr, stream))

    def integrate_device(self, n, dt_ptr, y_in_ptr, y_out_ptr, status_ptr=None, q_ptr=None, counters_ptr=None,
                         stream=None, order_ptr=None, rho_ptr=None, t0=0.0, pressure=None):
        """Device pointers (ints), enqueued on `stream` (int handle or None); `order_ptr`: visiting order;
        `rho_ptr`: per-cell densities of a constant-volume reactor; t0, pressure as in integrate()."""
        t0, P = float(t0), self._p(pressure)
        if rho_ptr:
            _check(lib().lsm_integrate_batch_device_rho(self.h, int(n), t0, P, dt_ptr, rho_ptr, y_in_ptr, y_out_ptr,
                                                        status_ptr, q_ptr, counters_ptr, order_ptr, stream))
        elif order_ptr:
            _check(lib().lsm_integrate_batch_device_ordered(self.h, int(n), t0, P, dt_ptr, y_in_ptr, y_out_ptr,
                                                            status_ptr, q_ptr, counters_ptr, order_ptr, stream))
        else:
            _check(lib().lsm_integrate_batch_device(self.h, int(n), t0, P, dt_ptr, y_in_ptr, y_out_ptr, status_ptr,
                                                    q_ptr, counters_ptr, stream))

    def integrate_device_noalloc(self, n, dt_ptr, y_in_ptr, y_out_ptr, status_ptr=None, q_ptr=None,
                                 counters_ptr=None, stream=None, order_ptr=None, t0=0.0, pressure=None):
        """integrate_device without any device allocation: the kernels' scratch is the library's
        static per-GPU array (lsm_integrate_batch_device_noalloc); calls on one GPU must not overlap."""
        _check(lib().lsm_integrate_batch_device_noalloc(self.h, int(n), float(t0), self._p(pressure), dt_ptr,
                                                        y_in_ptr, y_out_ptr, status_ptr, q_ptr, counters_ptr,
                                                        order_ptr, stream))

    def device_bytes(self):
        """Device bytes the handle holds now (tables, workspaces, buffers grown on demand)."""
        b = ctypes.c_uint64()
        _check(lib().lsm_device_bytes(self.h, ctypes.byref(b)))
        return b.value

    def cost_order(self, n, counters_ptr, order_ptr, stream=None):
        _check(lib().lsm_cost_order(self.h, int(n), counters_ptr, order_ptr, stream))

    def autotune(self, n, dt_ptr, y_ptr, sample=0, stream=None, t0=0.0, pressure=None):
        out = (ctypes.c_int32 * 3)()
        _check(lib().lsm_autotune(self.h, int(n), float(t0), self._p(pressure), dt_ptr, y_ptr, int(sample), stream,
                                  out))
        return dict(waves_per_block=out[0], grid=out[1], best_us=out[2], cells_per_wave=self.cells_per_wave())

    def launch(self):
        out = (ctypes.c_int32 * 5)()
        _check(lib().lsm_get_launch(self.h, out))
        names = {v: k for k, v in METHODS.items()}
        return dict(method=names.get(out[0], out[0]), waves_per_block=out[1], grid=out[2],
                    jacobian="analytic" if out[3] == 1 else "numerical", refine=out[4],
                    cells_per_wave=self.cells_per_wave())


def counters_dict(cnt: np.ndarray) -> dict:
    return {name: cnt[:, i] for i, name in enumerate(COUNTER_NAMES)}


class Server:
    """Device-resident serving of chemistry steps from several CFD cases
    (include/lsm_serve.h): HBM-budget admission, overlapped cases, per-case order."""

    def __init__(self, hbm_budget_bytes, max_inflight=4):
        h = ctypes.c_void_p()
        _check(lib().lsm_server_create(int(hbm_budget_bytes), int(max_inflight), ctypes.byref(h)))
        self.s = h
        self.steps = {}      # case id -> ChemistryStep (kept alive)
        self.neq = {}

    def add_case(self, step: ChemistryStep, y0):
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        cid = ctypes.c_int32()
        _check(lib().lsm_server_add_case(self.s, step.h, y0.shape[1], y0.ctypes.data, ctypes.byref(cid)))
        self.steps[cid.value] = step
        self.neq[cid.value] = y0.shape
        return cid.value

    def submit(self, case_id, dt, y=None, t0=0.0):
        """Queues a chemistry step from t0 over dt; y (optional): new fields for the case (host array
        [ns+1, n])."""
        t = ctypes.c_uint64()
        if y is not None:
            y = np.ascontiguousarray(y, dtype=np.float64)
            if y.shape != self.neq[case_id]:
                raise ValueError(f"y has shape {y.shape}, case has {self.neq[case_id]}")
        _check(lib().lsm_server_submit(self.s, int(case_id), float(t0), float(dt),
                                       y.ctypes.data if y is not None else None,
                                       ctypes.byref(t)))
        return t.value

    def poll(self, capacity=256, timeline=False):
        """Finished steps: (ticket, latency s, device s), with timeline=True also the step's device
        start and end (s after the server's epoch event)."""
        tk = np.zeros(capacity, dtype=np.uint64)
        lat = np.zeros(capacity)
        dev = np.zeros(capacity)
        t0 = np.zeros(capacity)
        t1 = np.zeros(capacity)
        nd = ctypes.c_int32()
        _check(lib().lsm_server_poll(self.s, tk.ctypes.data, lat.ctypes.data, dev.ctypes.data, t0.ctypes.data,
                                     t1.ctypes.data, capacity, ctypes.byref(nd)))
        k = nd.value
        if timeline:
            return [(int(tk[i]), float(lat[i]), float(dev[i]), float(t0[i]), float(t1[i])) for i in range(k)]
        return [(int(tk[i]), float(lat[i]), float(dev[i])) for i in range(k)]

    def drain(self):
        _check(lib().lsm_server_drain(self.s))

    def state(self, case_id):
        y = np.empty(self.neq[case_id])
        _check(lib().lsm_server_read_state(self.s, int(case_id), y.ctypes.data))
        return y

    def stats(self):
        out = np.zeros(8, dtype=np.int64)
        _check(lib().lsm_server_stats(self.s, out.ctypes.data))
        return dict(state_bytes=int(out[0]), peak_bytes=int(out[1]), evictions=int(out[2]), admitted=int(out[3]),
                    finished=int(out[4]), cases=int(out[5]), handle_bytes=int(out[6]), budget=int(out[7]))

    def close(self):
        if self.s:
            lib().lsm_server_destroy(self.s)
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
