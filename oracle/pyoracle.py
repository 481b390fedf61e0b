"""ctypes front-end of the CPU oracle (liblsm_oracle.so).

Test infrastructure only: imported by tests/, the smoke check and the CPU
baseline of bench.py, never by the laminarsmoke_amd package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblsm_oracle.so")
NCOUNTERS = 12
_lib = None

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int32)
RHS_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, _d, _d)


class OracleOptions(ctypes.Structure):
    _fields_ = [
        ("rel_tol", ctypes.c_double), ("abs_tol", ctypes.c_double),
        ("full_pivoting", ctypes.c_int32), ("max_steps", ctypes.c_int32),
        ("pressure", ctypes.c_double), ("min_temperature", ctypes.c_double),
        ("energy_equation", ctypes.c_int32), ("mass_fractions_tol", ctypes.c_double),
        ("t0", ctypes.c_double), ("constant_volume", ctypes.c_int32), ("strang_compact", ctypes.c_int32),
        ("rho", _d), ("jac_increment_scale", ctypes.c_double), ("reverse_sums", ctypes.c_int32),
    ]


class BdfOptions(ctypes.Structure):
    _fields_ = [
        ("rel_tol", ctypes.c_double), ("abs_tol", ctypes.c_double),
        ("full_pivoting", ctypes.c_int32), ("max_steps", ctypes.c_int32),
        ("use_min", ctypes.c_int32), ("use_max", ctypes.c_int32),
        ("ymin", _d), ("ymax", _d),
    ]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-B" if force else "-s", "-C", HERE])
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.lsm_oracle_rhs.restype = ctypes.c_double
        L.lsm_oracle_rhs_cv.restype = ctypes.c_double
        L.lsm_oracle_bdf.restype = ctypes.c_int
        L.lsm_oracle_integrate_cell.restype = ctypes.c_int
        L.lsm_oracle_lu_factor.restype = ctypes.c_int
        L.lsm_oracle_chemistry_step.argtypes = [
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, _d, _d, _d, _i, _d, _i, ctypes.c_int]
        _lib = L
    return _lib


def _p(a, t=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(t)) if a is not None else None


def options(rtol=1e-7, atol=1e-12, full_pivoting=False, max_steps=0, pressure=101325.0,
            min_temperature=0.0, energy=True, mass_fractions_tol=1e-4, t0=0.0, constant_volume=False,
            compact=False, rho=None, jac_increment_scale=1.0, reverse_sums=False) -> OracleOptions:
    """rho: per-cell densities of the constant-volume reactor (kept alive by the caller);
    jac_increment_scale: factor on the FD-Jacobian increments (1: the reference's);
    reverse_sums: formation rates summed over reactions in reverse order (False: the reference's)."""
    return OracleOptions(rtol, atol, int(full_pivoting), int(max_steps), pressure, min_temperature,
                         int(energy), mass_fractions_tol, t0, int(constant_volume), int(compact), _p(rho),
                         float(jac_increment_scale), int(bool(reverse_sums)))


def species_thermo(mech, T):
    ns = mech.ns
    cp, h, s = np.zeros(ns), np.zeros(ns), np.zeros(ns)
    lib().lsm_oracle_species_thermo(ctypes.byref(mech.struct()), ctypes.c_double(T), _p(cp), _p(h), _p(s))
    return cp, h, s


def kinetics(mech, T, c, P=101325.0):
    """Kinetics at (T, c) with the map at pressure P (it enters PLOG constants only)."""
    c = np.ascontiguousarray(c, dtype=np.float64)
    nr, ns = mech.nr, mech.ns
    kf, uK, km, r, R = np.zeros(nr), np.zeros(nr), np.zeros(nr), np.zeros(nr), np.zeros(ns)
    lib().lsm_oracle_kinetics_p(ctypes.byref(mech.struct()), ctypes.c_double(T), ctypes.c_double(P), _p(c),
                                _p(kf), _p(uK), _p(km), _p(r), _p(R))
    return dict(kf=kf, uKeq=uK, kmod=km, r=r, R=R)


def transport(tr, T, P, x):
    """mu, lambda, dmix[ns], teta[ns] of one cell (x: mole fractions)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    ns = tr.ns
    mu, lam = ctypes.c_double(), ctypes.c_double()
    dmix, teta = np.zeros(ns), np.zeros(ns)
    lib().lsm_oracle_transport(ctypes.byref(tr.struct()), ctypes.c_double(T), ctypes.c_double(P), _p(x),
                               ctypes.byref(mu), ctypes.byref(lam), _p(dmix), _p(teta))
    return mu.value, lam.value, dmix, teta


def refresh_cell(mech, tr, opts, P, y):
    """mu, lambda, dmix[ns], teta-or-Dsoret[ns] of one cell y[ns+1] with the refresh options."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    ns = tr.ns
    mu, lam = ctypes.c_double(), ctypes.c_double()
    dmix, teta = np.zeros(ns), np.zeros(ns)
    lib().lsm_oracle_refresh_cell(ctypes.byref(mech.struct()), ctypes.byref(tr.struct()),
                                  ctypes.byref(opts.struct()) if opts is not None else None, ctypes.c_double(P),
                                  _p(y), ctypes.byref(mu), ctypes.byref(lam), _p(dmix), _p(teta))
    return mu.value, lam.value, dmix, teta


def rhs(mech, y, P=101325.0, energy=True, rho0=None):
    """Constant-pressure RHS, or the constant-volume RHS at density rho0 when given."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    f = np.zeros_like(y)
    if rho0 is None:
        qr = lib().lsm_oracle_rhs(ctypes.byref(mech.struct()), ctypes.c_double(P), int(energy), _p(y), _p(f))
    else:
        qr = lib().lsm_oracle_rhs_cv(ctypes.byref(mech.struct()), ctypes.c_double(rho0), int(energy), _p(y), _p(f))
    return f, qr


def lu_factor(A, full=False):
    n = A.shape[0]
    LU = np.ascontiguousarray(A, dtype=np.float64).copy()
    rp = np.zeros(n, dtype=np.int32)
    cp = np.zeros(n, dtype=np.int32)
    lib().lsm_oracle_lu_factor(n, _p(LU), _p(rp, ctypes.c_int32), _p(cp, ctypes.c_int32), int(full))
    return LU, rp, cp


def lu_solve(LU, rp, cp, b, full=False):
    x = np.ascontiguousarray(b, dtype=np.float64).copy()
    lib().lsm_oracle_lu_solve(LU.shape[0], _p(LU), _p(rp, ctypes.c_int32), _p(cp, ctypes.c_int32), _p(x), int(full))
    return x


def bdf(f, y0, t0, tout, rtol=1e-7, atol=1e-12, ymin=None, ymax=None, full_pivoting=False, max_steps=0):
    """Gear/BDF on a Python callable f(y) -> dy/dt (small problems only)."""
    y0 = np.ascontiguousarray(y0, dtype=np.float64)
    n = y0.size

    def cb(_ctx, yp, fp):
        y = np.ctypeslib.as_array(yp, shape=(n,))
        out = np.ctypeslib.as_array(fp, shape=(n,))
        out[:] = f(y.copy())

    cfun = RHS_FN(cb)
    lo = np.ascontiguousarray(ymin, dtype=np.float64) if ymin is not None else None
    hi = np.ascontiguousarray(ymax, dtype=np.float64) if ymax is not None else None
    o = BdfOptions(rtol, atol, int(full_pivoting), int(max_steps), int(lo is not None), int(hi is not None),
                   _p(lo), _p(hi))
    yout = np.zeros(n)
    cnt = np.zeros(NCOUNTERS, dtype=np.int32)
    st = lib().lsm_oracle_bdf(n, cfun, None, ctypes.byref(o), ctypes.c_double(t0), _p(y0),
                              ctypes.c_double(tout), _p(yout), _p(cnt, ctypes.c_int32))
    return st, yout, cnt


def integrate_cell(mech, y, dt, **kw):
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros_like(y)
    q = ctypes.c_double(0.0)
    cnt = np.zeros(NCOUNTERS, dtype=np.int32)
    o = options(**kw)
    st = lib().lsm_oracle_integrate_cell(ctypes.byref(mech.struct()), ctypes.byref(o), ctypes.c_double(dt),
                                         _p(y), _p(out), ctypes.byref(q), _p(cnt, ctypes.c_int32))
    return st, out, q.value, cnt


def chemistry_step(mech, y_soa, dt, threads=1, q=None, rho=None, **kw):
    """SoA y[neq, n] -> (y_out, status, q, counters[n, 12]).  rho: per-cell densities of the
    constant-volume reactor (constant_volume=True), default the ideal-gas density of the input."""
    y_soa = np.ascontiguousarray(y_soa, dtype=np.float64)
    neq, n = y_soa.shape
    if rho is not None:
        rho = np.ascontiguousarray(np.broadcast_to(np.asarray(rho, dtype=np.float64), (n,)))
        kw["rho"] = rho
    dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (n,)))
    yout = np.zeros_like(y_soa)
    st = np.zeros(n, dtype=np.int32)
    qa = np.zeros(n) if q is None else np.ascontiguousarray(q, dtype=np.float64).copy()
    cnt = np.zeros((n, NCOUNTERS), dtype=np.int32)
    o = options(**kw)
    lib().lsm_oracle_chemistry_step(ctypes.byref(mech.struct()), ctypes.byref(o), n, _p(dt), _p(y_soa),
                                    _p(yout), _p(st, ctypes.c_int32), _p(qa), _p(cnt, ctypes.c_int32),
                                    int(threads))
    return yout, st, qa, cnt
