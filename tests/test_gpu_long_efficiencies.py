"""Third-body efficiency lists longer than the kinetics kernel's first block
of 8 (the rest of each list is summed in groups of 4): gri30 with every
third-body reaction given 9-22 efficiencies (conftest.extend_efficiencies;
the oracle's handling of the file is pinned by test_oracle_efficiencies.py).
GPU against the oracle on the same compiled tables: rates to the 1e-12
normalised bar of test_gpu_rates.py, Jacobians to the bars of
test_gpu_jacobian.py, reactor trajectories to the same-algorithm bar of
test_gpu_reactor.py (|dT|/T <= 1e-6, |dY| <= 1e-9 + 1e-6 |Y|).
"""
import numpy as np
import pytest

from conftest import extend_efficiencies, mech_file, parse_stoich

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def long_eff(tmp_path_factory):
    dst = str(tmp_path_factory.mktemp("mech") / "gri30_long_eff.mech")
    extend_efficiencies(mech_file("gri30"), dst)
    return dst


def test_rates_long_efficiency_lists(long_eff, oracle_mod):
    import cantera_amd as ca
    from cantera_amd.workloads import MechInfo
    from test_gpu_rates import TOL, random_states
    info = MechInfo(long_eff)
    om = oracle_mod.Mech(long_eff)
    sol = ca.Solution(long_eff)
    nu_r, nu_p = parse_stoich(long_eff)
    absnu = nu_r + nu_p
    for n, seed in ((1, 1), (65, 3), (300, 4)):
        T, P, Y = random_states(info, n, seed)
        w = sol.net_production_rates(T, P, Y.T)
        assert np.all(np.isfinite(w))
        worst = 0.0
        for s in range(n):
            r = om.rates(T[s], P[s], Y[s])
            scale = absnu @ (np.abs(r["ropf"]) + np.abs(r["ropr"])) + 1e-300
            worst = max(worst, float((np.abs(w[:, s] - r["wdot"]) / scale).max()))
        print(f"long efficiency lists n={n}: max normalised error {worst:.3e}", flush=True)
        assert worst <= TOL


def test_jacobian_long_efficiency_lists(long_eff, oracle_mod):
    import cantera_amd as ca
    from cantera_amd.workloads import MechInfo
    from test_gpu_jacobian import check_against_oracle, states
    info = MechInfo(long_eff)
    om = oracle_mod.Mech(long_eff)
    sol = ca.Solution(long_eff)
    T, P, Y = states(info, 40, 11)
    J = sol.reactor_jacobians(T, P, Y.T)
    assert np.all(np.isfinite(J))
    ec, et = check_against_oracle(om, T, P, Y, J)
    print(f"long efficiency lists: species columns {ec:.2e}, T column {et:.2e}", flush=True)
    assert ec <= 1e-10
    assert et <= 1e-5


def test_reactors_long_efficiency_lists(long_eff, oracle_mod):
    import cantera_amd as ca
    from cantera_amd.workloads import MechInfo, config3
    sol = ca.Solution(long_eff)
    om = oracle_mod.Mech(long_eff)
    wl = config3(MechInfo(long_eff), n=32, nsteps=8)
    n, nout = len(wl.T0), 8
    rb = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol)
    gpu = np.zeros((n, nout + 1, sol.n_species + 2))
    gpu[:, 0] = rb.state
    for k in range(1, nout + 1):
        rb.advance(k * wl.dt)
        gpu[:, k] = rb.state
    assert np.all(rb.status == 0)
    ref, _ = om.run_reactors(wl.T0, wl.P, wl.Y0, wl.dt, nout, wl.rtol, wl.atol, analytic=True, nthreads=8)
    dT = np.abs(gpu[:, :, 1] - ref[:, :, 1]) / ref[:, :, 1]
    dY = np.abs(gpu[:, :, 2:] - ref[:, :, 2:])
    print(f"long efficiency lists reactors: max dT/T {dT.max():.3e}, max dY {dY.max():.3e}", flush=True)
    assert dT.max() <= 1e-6
    assert (dY <= 1e-9 + 1e-6 * np.abs(ref[:, :, 2:])).all()
