"""Mechanisms beyond the dense integrators' 126 species on the GPU: the
synthetic 153-species GRI-3.0 + n-dodecane mechanism of
tests/golden/gri30_ndodecane.mech (pinned in test_oracle_large_mech.py).

* net production rates (rate kernel with four state slots per lane) against
  the oracle, the rates tests' normalised bar (1e-12 of the cancelling terms);
* mole-based reactor batches (preconditioned GMRES, four slots per lane),
  with third-body and falloff derivatives skipped and with the default
  derivatives, against the oracle's same algorithm: up to the ignition of
  the n-dodecane block with the bars of test_gpu_mole.py (|dT| <= 1e-6 T,
  |dY| <= 1e-9 + 1e-6 |Y|); after it, where the two runs' rounding
  differences have passed through a thermal runaway, T still at 1e-6 and Y
  with the reference's bars for two integrations of one reactor
  (test/python/test_reactor.py:1696-1728: rel 5e-4 + 1e-6);
* the entry points that keep the 126-species limit refuse the mechanism with
  a clear error.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, parse_stoich

pytestmark = pytest.mark.gpu
MERGED = os.path.join(GOLDEN, "gri30_ndodecane.mech")
ONE_ATM = 101325.0


def _mixture(sol):
    X = np.zeros(sol.n_species)
    for name, v in (("CH4", 0.5), ("O2", 2.0), ("N2", 7.52), ("c12h26", 0.05), ("o2", 0.93), ("n2", 3.5)):
        X[sol.species_index(name)] = v
    return sol.mass_fractions_from_moles(X / X.sum())


def test_rates_of_the_merged_mechanism(oracle_mod):
    import cantera_amd as ca
    from cantera_amd.workloads import MechInfo
    info = MechInfo(MERGED)
    om = oracle_mod.Mech(MERGED)
    sol = ca.Solution(MERGED)
    assert sol.n_species == 153
    nu_r, nu_p = parse_stoich(MERGED)
    absnu = nu_r + nu_p
    g = np.random.default_rng(21)
    n = 130
    T = g.uniform(300.0, 2500.0, n)
    P = ONE_ATM * np.exp(g.uniform(np.log(0.1), np.log(50.0), n))
    X = 10.0 ** g.uniform(-12, 0, (n, om.K))
    X /= X.sum(axis=1, keepdims=True)
    Y = info.mass_fractions_from_moles(X)
    w = sol.net_production_rates(T, P, Y.T)
    worst = 0.0
    for s in range(n):
        r = om.rates(T[s], P[s], Y[s])
        scale = absnu @ (np.abs(r["ropf"]) + np.abs(r["ropr"])) + 1e-300
        worst = max(worst, float((np.abs(w[:, s] - r["wdot"]) / scale).max()))
    print(f"merged mechanism rates: worst normalised error {worst:.2e}", flush=True)
    assert worst <= 1e-12


@pytest.mark.parametrize("skip", [True, False])
def test_mole_batch_on_the_merged_mechanism(oracle_mod, skip):
    import cantera_amd as ca
    sol = ca.Solution(MERGED)
    Y = _mixture(sol)
    T0 = np.array([950.0, 1000.0, 1050.0])
    n = len(T0)
    P = 20 * ONE_ATM
    mb = ca.MoleReactorBatch(sol, T0, P, np.tile(Y, (n, 1)), rtol=1e-8, atol=1e-16, skip_third_bodies=skip,
                             skip_falloff=skip, user_atol=True)
    om = oracle_mod.Mech(MERGED)
    refs = [oracle_mod.MoleReactor(om, T0[s], P, Y, rtol=1e-8, atol=1e-16, skip_third_bodies=skip,
                                   skip_falloff=skip, user_atol=True) for s in range(n)]
    worst_T = worst_Y = 0.0
    for t in (1e-4, 1e-3, 3e-3, 6e-3):
        print(f"merged skip={skip} t={t:g}", flush=True)
        mb.advance(np.full(n, t))
        assert np.all(mb.status == 0)
        Tm, Ym = mb.T, mb.Y
        for s in range(n):
            refs[s].advance(t)
            yo = refs[s].state
            Yo = refs[s].mass_fractions()
            dT = abs(Tm[s] - yo[0]) / yo[0]
            assert dT <= 1e-6, (t, s, Tm[s], yo[0])
            if t <= 3e-3:
                dY = float(np.max(np.abs(Ym[s] - Yo) / (1e-9 + 1e-6 * np.abs(Yo))))
                worst_T, worst_Y = max(worst_T, dT / 1e-6), max(worst_Y, dY)
                assert dY <= 1.0, (t, s)
            else:
                assert Ym[s] == pytest.approx(Yo, rel=5e-4, abs=1e-6)
    assert np.all(mb.T > 1500.0)        # both blocks ignited
    ls = mb.linear_stats
    assert np.all(ls["lin_iters"] > 0) and np.all(ls["prec_evals"] == mb.stats["jac"])
    print(f"merged skip={skip}: worst T {worst_T:.3f}, Y {worst_Y:.3f} of the bars up to 3 ms; GPU steps "
          f"{mb.stats['steps'].tolist()}, oracle {refs[0].stats['steps']}", flush=True)


def test_dense_entry_points_refuse_more_than_126_species():
    import cantera_amd as ca
    sol = ca.Solution(MERGED)
    Y = _mixture(sol).reshape(1, -1)
    with pytest.raises(ca.CanteraGpuError, match="126 species"):
        ca.ReactorBatch(sol, [1000.0], ONE_ATM, Y)
    with pytest.raises(ca.CanteraGpuError, match="126 species"):
        sol.reactor_jacobians(np.array([1000.0]), np.array([ONE_ATM]), Y.T)
    with pytest.raises(ca.CanteraGpuError, match="126 species"):
        sol.equilibrate(np.array([1000.0]), np.array([ONE_ATM]), Y)
