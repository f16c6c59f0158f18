"""Golden fixture: h2o2 constant-pressure reactor trajectories from scipy's
Radau IIA (an implicit Runge-Kutta method, independent of the BDF under test)
at rtol 1e-12 on the oracle's right-hand side.

Run from the repository root:  python tests/golden/make_radau_h2o2.py
Writes tests/golden/radau_h2o2.npz (inputs and states at 20 output times).
"""
import os
import sys

import numpy as np
from scipy.integrate import solve_ivp

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import Mech  # noqa: E402
from cantera_amd.workloads import MechInfo, config1  # noqa: E402


def main():
    path = os.path.join(ROOT, "cantera_amd", "mechanisms", "h2o2.mech")
    om = Mech(path)
    wl = config1(MechInfo(path), n=1024)
    pick = [0, 1, 2, 3, 5, 8]
    T0, P, Y0 = wl.T0[pick], wl.P[pick], wl.Y0[pick]
    touts = np.linspace(0.0, 1e-3, 21)
    states = np.zeros((len(pick), len(touts), om.K + 2))
    for i in range(len(pick)):
        rho = P[i] / (8314.46261815324 * T0[i]) / np.sum(Y0[i] / om.mw)
        y0 = np.concatenate([[rho], [T0[i]], Y0[i]])
        sol = solve_ivp(lambda t, y: om.rhs(P[i], y), (0.0, touts[-1]), y0, method="Radau",
                        t_eval=touts, rtol=1e-12, atol=1e-20, jac=lambda t, y: om.jacobian(P[i], y, True))
        assert sol.success, sol.message
        states[i] = sol.y.T
    np.savez(os.path.join(ROOT, "tests", "golden", "radau_h2o2.npz"), T0=T0, P=P, Y0=Y0, touts=touts, states=states)
    print("wrote radau_h2o2.npz", states.shape)


if __name__ == "__main__":
    main()
