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


_STOICH_O2 = {"H2": 0.5, "CO": 0.5, "CH4": 2.0, "CH2O": 1.0, "C2H4": 3.0, "C2H6": 3.5}
_DILUENTS = ("N2", "AR", "HE")


def premixed_cells(mech: Mechanism, n: int, seed: int, T_lo: float = 900.0, T_hi: float = 2200.0,
                   radical_fraction: float = 0.5):
    """Fuel/air cells for a real (CHEMKIN) mechanism.

    The fuels of the mechanism among H2, CO, CH4, CH2O, C2H4, C2H6 in random
    proportions (Dirichlet), O2 at an equivalence ratio log-uniform in [0.3, 3],
    3.76 mol of diluent per mol of O2 (N2, with a random share of AR or HE where
    the mechanism has them).  In ``radical_fraction`` of the cells every other
    species (radicals, intermediates, products) gets a mass fraction log-uniform in
    [1e-8, 1e-3], as in a flame front.  T ~ U[T_lo, T_hi] K.
    """
    rng = np.random.default_rng(seed)
    sp = mech.species
    fuels = [s for s in _STOICH_O2 if s in sp]
    dil = [s for s in _DILUENTS if s in sp]
    if not fuels or "O2" not in sp or not dil:
        raise ValueError("premixed_cells needs a fuel, O2 and a diluent in the mechanism")
    x = np.zeros((mech.ns, n))
    share = rng.dirichlet(np.ones(len(fuels)), size=n).T if len(fuels) > 1 else np.ones((1, n))
    phi = np.exp(rng.uniform(np.log(0.3), np.log(3.0), n))
    o2 = sum(share[i] * _STOICH_O2[f] for i, f in enumerate(fuels)) / phi
    for i, f in enumerate(fuels):
        x[sp.index(f)] = share[i]
    x[sp.index("O2")] = o2
    d = rng.dirichlet(np.array([4.0] + [1.0] * (len(dil) - 1)), size=n).T if len(dil) > 1 else np.ones((1, n))
    for i, s in enumerate(dil):
        x[sp.index(s)] += 3.76 * o2 * d[i]
    y = x * mech.mw[:, None]
    y /= y.sum(axis=0, keepdims=True)
    others = [k for k, s in enumerate(sp) if s not in fuels and s != "O2" and s not in dil]
    hot = rng.random(n) < radical_fraction
    y[np.ix_(others, np.where(hot)[0])] = np.exp(rng.uniform(np.log(1e-8), np.log(1e-3),
                                                            size=(len(others), int(hot.sum()))))
    y /= y.sum(axis=0, keepdims=True)
    out = np.empty((mech.neq, n))
    out[:mech.ns] = y
    out[mech.ns] = rng.uniform(T_lo, T_hi, n)
    return out


def coflow_premixed_field(mech: Mechanism, nx: int, ny: int, seed: int, flame_fraction: float = 0.4):
    """The coflow geometry of :func:`coflow_field` for a real (CHEMKIN) mechanism: the
    flame band holds :func:`premixed_cells` at 1200-2200 K with a radical pool, the rest
    fresh fuel/air at 300-600 K."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    z = (np.arange(ny) + 0.5) / ny
    X, Z = np.meshgrid(x, z, indexing="ij")
    r_f = 0.12 + 0.25 * np.sqrt(Z)
    phi = -np.abs(X - r_f) + 0.02 * rng.standard_normal(X.shape)
    flame = (phi >= np.quantile(phi, 1.0 - flame_fraction)).ravel()
    n = nx * ny
    y = np.empty((mech.neq, n))
    nf = int(flame.sum())
    y[:, flame] = premixed_cells(mech, nf, seed + 1, 1200.0, 2200.0, radical_fraction=1.0)
    y[:, ~flame] = premixed_cells(mech, n - nf, seed + 2, 300.0, 600.0, radical_fraction=0.0)
    return y, flame


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
