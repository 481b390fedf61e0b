"""Seeded synthetic cell fields (mass fractions + temperature) for the workloads.

Fields are returned in the solver layout: ``y[neq, n_cells]`` (species rows, then
temperature), the same species-major SoA as OpenFOAM's ``Y[i][celli]`` fields read
by the chemistry loop (LS/unsteady/chemistry_DI.H:101-104).
"""
from __future__ import annotations

import numpy as np

from .mechanism import Mechanism, workload_mechanism

P_ATM = 101325.0


def is_radical(name: str) -> bool:
    from .mechanism import _parse
    a = _parse(name)
    return bool(int(a[1]) % 2 == 1 or a.sum() == 1)


_STABLE_ALPHA = {"N2": 8.0, "O2": 2.0, "H2": 1.0, "CH4": 1.0, "H2O": 0.5, "CO2": 0.5, "CO": 0.3}


def _stable_mix(mech: Mechanism, n: int, rng) -> np.ndarray:
    idx = [mech.species.index(s) for s in _STABLE_ALPHA if s in mech.species]
    alpha = np.array([_STABLE_ALPHA[mech.species[i]] for i in idx])
    y = np.zeros((mech.ns, n))
    y[idx, :] = rng.dirichlet(alpha, size=n).T
    return y


def dirichlet_cells(mech: Mechanism, n: int, T_lo: float, T_hi: float, seed: int, alpha: float = 0.5):
    """Config-1 style cells: T ~ U[T_lo, T_hi], Y ~ Dirichlet(alpha * 1)."""
    rng = np.random.default_rng(seed)
    y = np.empty((mech.neq, n))
    y[:mech.ns] = rng.dirichlet(np.full(mech.ns, alpha), size=n).T
    y[mech.ns] = rng.uniform(T_lo, T_hi, n)
    return y


def coflow_field(mech: Mechanism, nx: int, ny: int, seed: int, flame_fraction: float = 0.4):
    """2-D coflow-like field: a flame sheet band (flame_fraction of the cells) at
    1200-2200 K with Dirichlet(0.5) composition, the rest cold (300-600 K) fuel/air
    mixture of stable species."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    z = (np.arange(ny) + 0.5) / ny
    X, Z = np.meshgrid(x, z, indexing="ij")
    r_f = 0.12 + 0.25 * np.sqrt(Z)                       # flame sheet radius grows downstream
    phi = -np.abs(X - r_f) + 0.02 * rng.standard_normal(X.shape)
    thr = np.quantile(phi, 1.0 - flame_fraction)
    flame = (phi >= thr).ravel()
    n = nx * ny
    y = np.zeros((mech.neq, n))
    y[:mech.ns] = _stable_mix(mech, n, rng)
    nf = int(flame.sum())
    y[:mech.ns, flame] = rng.dirichlet(np.full(mech.ns, 0.5), size=nf).T
    T = rng.uniform(300.0, 600.0, n)
    T[flame] = rng.uniform(1200.0, 2200.0, nf)
    y[mech.ns] = T
    return y, flame


def stiffness_spread_cells(mech: Mechanism, n: int, seed: int):
    """Config-3 cells: T ~ U[300, 2500] K, radicals log-uniform in [1e-12, 1e-3],
    stable species from a Dirichlet fuel/air mixture."""
    rng = np.random.default_rng(seed)
    y = _stable_mix(mech, n, rng)
    rad = [k for k, s in enumerate(mech.species) if is_radical(s)]
    y[rad, :] = np.exp(rng.uniform(np.log(1e-12), np.log(1e-3), size=(len(rad), n)))
    y /= y.sum(axis=0, keepdims=True)
    out = np.empty((mech.neq, n))
    out[:mech.ns] = y
    out[mech.ns] = rng.uniform(300.0, 2500.0, n)
    return out


def hot_cells(mech: Mechanism, n: int, seed: int, T_lo=800.0, T_hi=2400.0):
    """Config-4 cells: T ~ U[800, 2400] K, stable mixture with log-uniform radicals."""
    rng = np.random.default_rng(seed)
    y = stiffness_spread_cells(mech, n, seed)
    y[mech.ns] = rng.uniform(T_lo, T_hi, n)
    return y


# The five workloads (cells are generated lazily; sizes are the full sizes).
WORKLOADS = {
    1: dict(name="cfg1-14sp-4096", ns=14, cells=4096, dt=1e-5, rtol=1e-7, atol=1e-12),
    2: dict(name="cfg2-53sp-coflow-512x512", ns=53, cells=512 * 512, dt=1e-5, rtol=1e-7, atol=1e-12),
    3: dict(name="cfg3-53sp-spread-1M", ns=53, cells=1 << 20, dt=1e-5, rtol=1e-7, atol=1e-12),
    4: dict(name="cfg4-160sp-131072", ns=160, cells=131072, dt=1e-5, rtol=1e-6, atol=1e-12),
}


def workload_cells(cfg: int, n: int | None = None, seed: int | None = None):
    """Cells of workload ``cfg`` (first ``n`` cells of the seeded field if given)."""
    w = WORKLOADS[cfg]
    mech = workload_mechanism(w["ns"])
    n_full = w["cells"]
    seed = {1: 11, 2: 22, 3: 33, 4: 44}[cfg] if seed is None else seed
    if cfg == 1:
        y = dirichlet_cells(mech, n_full, 900.0, 2400.0, seed)
    elif cfg == 2:
        y, _ = coflow_field(mech, 512, 512, seed)
    elif cfg == 3:
        y = stiffness_spread_cells(mech, n_full, seed)
    else:
        y = hot_cells(mech, n_full, seed)
    if n is not None and n < y.shape[1]:
        # seeded subsample that keeps the field's mix (stratified by stride)
        idx = np.linspace(0, y.shape[1] - 1, n).astype(np.int64)
        y = y[:, idx]
    return mech, np.ascontiguousarray(y)
