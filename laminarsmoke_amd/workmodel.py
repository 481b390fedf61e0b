"""Algorithmic work model of the chemistry step (FP64 operations per cell).

Counts follow the per-operation costs fixed from the mechanism tables:

* F_rhs  = 5 Ns (mass->mole, concentrations) + 28 Ns (NASA cp, h, s) + 4 Ns (cp_mix, QR)
           + 2 Ns (dY, dT) + sum_j [4 (Arrhenius) + 4 nnz_j (dH, dS) + 4 rev_j (1/Kc)
           + nnz_j (prod c) + 3 (net) + nnz_j (formation)] + sum_tb (2 n_eff + 1) + 25 N_troe,
           each exp/log/log10/pow counted as one operation (and reported as special_ops);
* F_jac  = N^2 on top of the N RHS evaluations of the forward-difference Jacobian;
  analytic: F_jac_an = F_rhs (rates at the state) + sum_j 2 nnz_j (nnz_j + 2) (dr_j/dc of
  the participants, scattered to every participant's row) + sum_{tb, fo} 2 nnz_j (Ns + n_eff)
  (collider columns) + 25 per Troe (dF/dT, dF/dPr) + 8 N^2 (S_k, column sums, assembly of
  the N x N block);
* F_fact = (2/3) N^3 + N^2 (Newton matrix G = I - h l0 J and its LU);
* F_solve = 2 N^2; F_step = 27 N (Nordsieck prediction/correction, norms at order ~3).

W_cell = n_rhs F_rhs + n_jac F_jac + n_fact F_fact + n_solve F_solve + n_steps F_step, with
the n_* taken from the per-cell counters the kernels write.
"""
from __future__ import annotations

import numpy as np


def per_op_costs(mech) -> dict:
    ns, nr = mech.ns, mech.nr
    N = ns + 1
    nnz = (np.diff(mech.reac_ptr) + np.diff(mech.prod_ptr)).astype(np.float64)
    rev = (mech.rflags & 1).astype(np.float64)
    n_eff_tb = np.diff(mech.tb_ptr).astype(np.float64)
    n_troe = int(np.sum(mech.fo_type == 1))
    f_rhs = (5 + 28 + 4 + 2) * ns + float(np.sum(4 + 4 * nnz + 4 * rev + nnz + 3 + nnz)) \
        + float(np.sum(2 * n_eff_tb + 1)) + 25 * n_troe
    special = nr + float(rev.sum()) + len(mech.fo_rxn) + 8 * n_troe + 2   # exp/log per RHS
    coll = np.zeros(nr)
    coll[mech.tb_rxn] = ns + n_eff_tb
    coll[mech.fo_rxn] = ns + np.diff(mech.fo_ptr)
    f_jac_an = f_rhs + float(np.sum(2 * nnz * (nnz + 2))) + float(np.sum(2 * nnz * coll)) + 25 * n_troe + 8.0 * N * N
    return dict(N=N, F_rhs=f_rhs, special_rhs=special, F_jac=float(N * N), F_jac_an=f_jac_an,
                F_fact=2.0 / 3.0 * N ** 3 + N * N, F_solve=2.0 * N * N, F_step=27.0 * N)


def flops_from_counters(mech, counters: np.ndarray, jacobian: str = "numerical") -> dict:
    """Total algorithmic FP64 operations of a batch from its [n, 12] counters.  With the
    analytic Jacobian the counters hold no Jacobian RHS evaluations and each Jacobian
    costs F_jac_an (including its own rate evaluation)."""
    c = per_op_costs(mech)
    cnt = counters.astype(np.float64)
    n_steps, n_jac, n_fact, n_solve, n_rhs = cnt[:, 0], cnt[:, 3], cnt[:, 4], cnt[:, 5], cnt[:, 9]
    fj = c["F_jac_an"] if jacobian == "analytic" else c["F_jac"]
    w = n_rhs * c["F_rhs"] + n_jac * fj + n_fact * c["F_fact"] + n_solve * c["F_solve"] + n_steps * c["F_step"]
    n_special = n_rhs + (n_jac if jacobian == "analytic" else 0.0)
    return dict(flops=float(w.sum()), special_ops=float((n_special * c["special_rhs"]).sum()),
                rhs=float(n_rhs.sum()), jacobians=float(n_jac.sum()), factorizations=float(n_fact.sum()),
                solves=float(n_solve.sum()), steps=float(n_steps.sum()), costs=c)


def hbm_bytes_min(mech, n_cells: int) -> float:
    """Minimal HBM traffic of one batch: state in and out (N doubles each way) plus dt,
    status, Q and counters."""
    N = mech.ns + 1
    return float(n_cells) * (2 * N * 8 + 8 + 4 + 8 + 12 * 4)
