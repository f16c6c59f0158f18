"""Batched chemical equilibrium on the GPU (cg_equilibrate_batch).

Bars:
* the reference's HP-equilibrium fixture (test/python/test_equilibrium.py:
  477-502, gas rows of gas-carbon-equil.csv / -5atm.csv): T and every
  species' kmol within the reference's rtol of 1e-7 (+1e-15 kmol absolute for
  species below that);
* against the oracle (same method on the CPU): T within 1e-10 relative, mass
  fractions within 1e-9 relative + 1e-18 absolute -- both solve the same
  Newton system to corrections below 1e-12, only summation order differs;
* element conservation to 1e-12 of each element's total;
* cxx_demo (h2o2, HP from 1200 K): equilibrium temperature to the blessed 5
  significant figures and detailed balance of the GPU rates there.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, mech_file

pytestmark = pytest.mark.gpu
ONE_ATM = 101325.0
R = 6.02214076e26 * 1.380649e-23


@pytest.mark.parametrize("fname,P", [("gas-carbon-equil.csv", ONE_ATM), ("gas-carbon-equil-5atm.csv", 5 * ONE_ATM)])
def test_hp_fixture_on_gpu(fname, P, oracle_mod):
    import cantera_amd as ca
    from cantera_amd.workloads import MechInfo, equivalence_ratio_moles
    info = MechInfo(mech_file("gri30"))
    sol = ca.Solution("gri30")
    rows = [r for r in np.loadtxt(os.path.join(GOLDEN, fname)) if r[-1] == 0.0]
    X = equivalence_ratio_moles(info, "CH4", [r[0] for r in rows])
    Y = info.mass_fractions_from_moles(X)
    n = len(rows)
    T, Ye, it = sol.equilibrate(np.full(n, 300.0), P, Y.T)
    print(f"{fname}: iterations {it.tolist()}", flush=True)
    om = oracle_mod.Mech(mech_file("gri30"))
    for i, row in enumerate(rows):
        W0 = 1.0 / np.sum(Y[i] / info.molecular_weights)
        kmol = Ye[:, i] / info.molecular_weights * W0
        ref = row[2:-1]
        assert T[i] == pytest.approx(row[1], rel=1e-7)
        assert np.all(np.abs(kmol - ref) <= 1e-7 * np.abs(ref) + 1e-15), row[0]
        To, Yo, _ = om.equilibrate(300.0, P, Y[i])
        assert abs(T[i] - To) <= 1e-10 * To
        assert np.all(np.abs(Ye[:, i] - Yo) <= 1e-9 * np.abs(Yo) + 1e-18)


def test_cxx_demo_equilibrium_and_detailed_balance():
    import cantera_amd as ca
    sol = ca.Solution("h2o2")
    X = np.zeros(sol.n_species)
    for name, v in (("H2", 1.0), ("O2", 1.0), ("AR", 2.0)):
        X[sol.species_index(name)] = v
    Y = sol.mass_fractions_from_moles(X / X.sum())
    T, Ye, _ = sol.equilibrate([1200.0], ONE_ATM, Y.reshape(-1, 1))
    assert T[0] == pytest.approx(2800.9, abs=0.05)
    w = sol.net_production_rates(T, [ONE_ATM], Ye)
    Wbar = 1.0 / np.sum(Ye[:, 0] / sol.molecular_weights)
    ctot = ONE_ATM / (R * T[0])
    # net rates vanish relative to the largest forward rate of progress there
    # (the blessed table: H2 + OH <=> H + H2O at 229.06 kmol/m3/s)
    assert np.max(np.abs(w)) <= 1e-8 * 229.06
    assert Wbar > 0 and ctot > 0


@pytest.mark.parametrize("mech,fuel,n", [("gri30", "CH4", 512), ("nDodecane", "c12h26", 64)])
def test_batch_hp_and_tp_against_oracle(mech, fuel, n, oracle_mod):
    """Random fuel/air states (phi 0.3-3, 1-40 atm, 300-1500 K) equilibrated
    HP and, at 1000-3000 K, TP: all converge, elements are conserved, and the
    oracle agrees state by state (a sample of 24)."""
    import cantera_amd as ca
    from cantera_amd.workloads import MechInfo, equivalence_ratio_moles
    info = MechInfo(mech_file(mech))
    sol = ca.Solution(mech)
    om = oracle_mod.Mech(mech_file(mech))
    g = np.random.default_rng(4)
    ox = (("O2", 1.0), ("N2", 3.76)) if mech == "gri30" else (("o2", 1.0), ("n2", 3.76))
    X = equivalence_ratio_moles(info, fuel, g.uniform(0.3, 3.0, n), oxidizer=ox)
    Y = info.mass_fractions_from_moles(X)
    P = ONE_ATM * np.exp(g.uniform(0.0, np.log(40.0), n))
    A = info.composition / info.molecular_weights[:, None]
    for mode, T0 in (("HP", g.uniform(300.0, 1500.0, n)), ("TP", g.uniform(1000.0, 3000.0, n))):
        T, Ye, it = sol.equilibrate(T0, P, Y.T, mode=mode)
        assert np.all(it > 0)
        b0, b1 = Y @ A, Ye.T @ A
        assert np.all(np.abs(b1 - b0) <= 1e-12 * b0.max(axis=1, keepdims=True))
        if mode == "TP":
            assert np.array_equal(T, T0)
        for s in np.linspace(0, n - 1, 24).astype(int):
            To, Yo, _ = om.equilibrate(T0[s], P[s], Y[s], mode=mode)
            assert abs(T[s] - To) <= 1e-10 * To, (mode, s)
            assert np.all(np.abs(Ye[:, s] - Yo) <= 1e-9 * np.abs(Yo) + 1e-18), (mode, s)
        print(f"{mech} {mode}: iterations median {np.median(it):.0f} max {it.max()}", flush=True)


def test_bad_mode_and_inputs():
    import cantera_amd as ca
    sol = ca.Solution("h2o2")
    Y = np.full((sol.n_species, 2), 1.0 / sol.n_species)
    with pytest.raises(ca.CanteraGpuError, match="HP"):
        sol.equilibrate([1000.0, 1000.0], ONE_ATM, Y, mode="UV")
    with pytest.raises(ca.CanteraGpuError, match="temperature"):
        sol.equilibrate([1000.0, -5.0], ONE_ATM, Y)
