"""Writes the golden vectors under tests/golden/ (run from the repository root:
``python tests/golden/make_golden.py``).

They come from SciPy's Radau IIA integrator (scipy.integrate.solve_ivp, method
"Radau"; SciPy 1.15.3), a third-party code that shares nothing with the oracle,
run at far tighter tolerances than the chemistry step uses:

* robertson.npz - Robertson's stiff kinetics (Hairer & Wanner, Solving ODEs II,
  sec. IV.1) at t = 0.4, 40, 4e3 s, rtol 1e-12;
* cfg1_truth.npz - 6 cells of the seeded 14-species workload over dt = 1e-5 s,
  integrated from the independent dense formulation (tests/dense_reference.py)
  at rtol 1e-11.
"""
import os
import sys

import numpy as np
from scipy.integrate import solve_ivp

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import dense_reference as dref  # noqa: E402
from laminarsmoke_amd.fields import workload_cells  # noqa: E402


def robertson(t, y):
    return [-0.04 * y[0] + 1e4 * y[1] * y[2], 0.04 * y[0] - 1e4 * y[1] * y[2] - 3e7 * y[1] ** 2, 3e7 * y[1] ** 2]


def main():
    ts = np.array([0.4, 40.0, 4.0e3])
    sol = solve_ivp(robertson, (0.0, ts[-1]), [1.0, 0.0, 0.0], method="Radau", rtol=1e-12, atol=1e-20,
                    t_eval=ts, dense_output=False)
    np.savez(os.path.join(HERE, "robertson.npz"), t=ts, y=sol.y.T)
    print("robertson", sol.y.T)

    mech, y = workload_cells(1, n=6)
    out = np.zeros_like(y)
    for i in range(y.shape[1]):
        y0 = y[:, i].copy()
        y0[:mech.ns] = np.clip(y0[:mech.ns], 0, None)
        y0[:mech.ns] /= y0[:mech.ns].sum()
        s = solve_ivp(lambda t, v: dref.rhs(mech, v)[0], (0.0, 1e-5), y0, method="Radau", rtol=1e-11,
                      atol=1e-20, first_step=1e-16)
        out[:, i] = s.y[:, -1]
        print(i, s.status, s.nfev, out[-1, i])
    np.savez(os.path.join(HERE, "cfg1_truth.npz"), y_in=y, y_out=out, dt=1e-5)


if __name__ == "__main__":
    main()
