"""Mole-based reactor batches with the sparse preconditioner and GMRES on the
GPU (cg_reactors_create_mole) against the oracle's same algorithm
(test_oracle_mole.py) and against the GPU's mass-fraction reactor.

Bars, per reactor and output:
* against the oracle (same method; they differ in the rounding of the RHS,
  the dot products and the LU order, which moves the step sequence at the
  level of the integration tolerance): as the plain integrator's parity tests,
  |dT| <= 1e-6 T, |dY| <= 1e-9 + 1e-6 |Y|;
* against the mass-fraction reactor (another state, another linear solver):
  the reference's own bars for that comparison (test/python/test_reactor.py:
  1696-1728: Y rel 5e-4 + 1e-6, T rel 1e-5);
* linear-solver counters keep CVODES' relations.
"""
import numpy as np
import pytest

from conftest import mech_file

pytestmark = pytest.mark.gpu
ONE_ATM = 101325.0

CASES = [
    ("h2o2", (("H2", 2.0), ("O2", 1.0), ("AR", 4.0)), [950.0, 1000.0, 1100.0, 1300.0], ONE_ATM,
     [1e-4, 2e-4, 5e-4, 1e-3]),
    ("gri30", (("CH4", 1.0), ("O2", 2.0), ("N2", 7.52)), [1150.0, 1200.0, 1400.0, 1800.0], ONE_ATM,
     [0.005, 0.05, 0.1, 0.2]),
    ("nDodecane", (("c12h26", 1.0), ("o2", 18.5), ("n2", 69.56)), [850.0, 900.0, 1100.0], 20 * ONE_ATM,
     [1e-4, 5e-4, 1e-3, 2e-3]),
]


def _Y(sol, comp):
    X = np.zeros(sol.n_species)
    for name, v in comp:
        X[sol.species_index(name)] = v
    return sol.mass_fractions_from_moles(X / X.sum())


@pytest.mark.parametrize("mech,comp,T0,P,times", CASES)
@pytest.mark.parametrize("skip", [False, True])
def test_mole_batch_matches_oracle_and_mass_fraction_batch(oracle_mod, mech, comp, T0, P, times, skip):
    import cantera_amd as ca
    sol = ca.Solution(mech)
    Y = _Y(sol, comp)
    n = len(T0)
    T0 = np.asarray(T0)
    mb = ca.MoleReactorBatch(sol, T0, P, np.tile(Y, (n, 1)), rtol=1e-8, atol=1e-16, skip_third_bodies=skip,
                             skip_falloff=skip, user_atol=True)
    yb = ca.ReactorBatch(sol, T0, P, np.tile(Y, (n, 1)), rtol=1e-8, atol=1e-16)
    om = oracle_mod.Mech(mech_file(mech))
    refs = [oracle_mod.MoleReactor(om, T0[s], P, Y, rtol=1e-8, atol=1e-16, skip_third_bodies=skip,
                                   skip_falloff=skip, user_atol=True) for s in range(n)]
    worst_T = worst_Y = 0.0
    for t in times:
        print(f"{mech} skip={skip} n={n} t={t:g}", flush=True)
        mb.advance(np.full(n, t))
        yb.advance(np.full(n, t))
        Tm, Ym = mb.T, mb.Y
        Tb, Yb = yb.T, yb.Y
        assert np.all(mb.status == 0)
        for s in range(n):
            refs[s].advance(t)
            yo = refs[s].state
            Yo = refs[s].mass_fractions()
            dT = abs(Tm[s] - yo[0]) / yo[0]
            dY = float(np.max(np.abs(Ym[s] - Yo) / (1e-9 + 1e-6 * np.abs(Yo))))
            worst_T, worst_Y = max(worst_T, dT / 1e-6), max(worst_Y, dY)
            assert dT <= 1e-6, (t, s, Tm[s], yo[0])
            assert dY <= 1.0, (t, s)
            assert Tm[s] == pytest.approx(Tb[s], rel=1e-5)
            assert Ym[s] == pytest.approx(Yb[s], rel=5e-4, abs=1e-6)
    st, ls = mb.stats, mb.linear_stats
    assert np.all(ls["lin_iters"] > 0) and np.all(ls["prec_evals"] == st["jac"])
    assert np.all(ls["jt_vec_prod_evals"] == ls["lin_iters"])
    assert np.all(ls["lin_rhs_evals"] == ls["lin_iters"])
    assert np.all(ls["prec_solves"] >= ls["lin_iters"])
    o = refs[0]
    print(f"{mech} skip={skip}: worst T {worst_T:.3f}, Y {worst_Y:.3f} of the bars; GPU steps {st['steps'].tolist()}, "
          f"oracle {o.stats['steps']}; Krylov iterations {ls['lin_iters'].tolist()}", flush=True)


def test_adaptive_precon_integration_constant_volume(oracle_mod):
    """test_adaptive_precon_integration (test/python/test_reactor.py:1696-1728)
    on the GPU: gri30 IdealGasMoleReactors with the preconditioner
    (skip-third-bodies, skip-falloff) at 900-1100 K, 1 atm, CH4/air phi = 1,
    atol 1e-16, rtol 1e-8, outputs every 0.5 s to 4.5 s; against the GPU's
    constant-volume mass-fraction reactors with the reference's bars (Y rel
    5e-4 + 1e-6, T and P rel 1e-5) and against the oracle's same algorithm
    (T 1e-6, Y 1e-9 + 1e-6 relative)."""
    import cantera_amd as ca
    from cantera_amd.workloads import equivalence_ratio_moles
    sol = ca.Solution("gri30")
    Y = sol.mass_fractions_from_moles(equivalence_ratio_moles(sol, "CH4", [1.0])[0])
    T0 = np.array([900.0, 950.0, 1000.0, 1100.0])
    n = len(T0)
    mb = ca.MoleReactorBatch(sol, T0, ONE_ATM, np.tile(Y, (n, 1)), rtol=1e-8, atol=1e-16, skip_third_bodies=True,
                             skip_falloff=True, user_atol=True, const_volume=True)
    db = ca.ReactorBatch(sol, T0, ONE_ATM, np.tile(Y, (n, 1)), rtol=1e-8, atol=1e-16, const_volume=True)
    om = oracle_mod.Mech(mech_file("gri30"))
    refs = [oracle_mod.MoleReactor(om, T0[s], ONE_ATM, Y, rtol=1e-8, atol=1e-16, skip_third_bodies=True,
                                   skip_falloff=True, user_atol=True, const_volume=True) for s in range(n)]
    R = 6.02214076e26 * 1.380649e-23
    for t in np.arange(0.5, 5.0, 0.5):
        print(f"gri30 IdealGasMoleReactor t={t:g}", flush=True)
        mb.advance(np.full(n, t))
        db.advance(np.full(n, t))
        ym, yd = mb.state, db.state
        assert np.all(ym[:, 1] == 1.0)
        for s in range(n):
            assert mb.Y[s] == pytest.approx(yd[s, 2:], rel=5e-4, abs=1e-6)
            assert ym[s, 0] == pytest.approx(yd[s, 1], rel=1e-5)
            Pm = R * ym[s, 0] * ym[s, 2:].sum() / ym[s, 1]
            Pd = yd[s, 0] * R * yd[s, 1] * np.sum(yd[s, 2:] / sol.molecular_weights)
            assert Pm == pytest.approx(Pd, rel=1e-5)
            refs[s].advance(t)
            yo = refs[s].state
            assert abs(ym[s, 0] - yo[0]) <= 1e-6 * yo[0]
            Yo = refs[s].mass_fractions()
            assert np.all(np.abs(mb.Y[s] - Yo) <= 1e-9 + 1e-6 * np.abs(Yo))
    assert np.all(mb.linear_stats["lin_iters"] > 0)


def test_mole_batch_rejections():
    import cantera_amd as ca
    sol = ca.Solution("h2o2")
    Y = _Y(sol, CASES[0][1])
    b = ca.MoleReactorBatch(sol, [1000.0], ONE_ATM, Y.reshape(1, -1))
    with pytest.raises(ca.CanteraGpuError, match="mole"):
        b.add_sensitivity_reaction(2)
    with pytest.raises(ca.CanteraGpuError, match="mass-fraction"):
        b.autotune([1e-6])
    from cantera_amd._lib import check
    with pytest.raises(ca.CanteraGpuError, match="flags"):
        check(ca.lib().cg_reactors_create_mole(sol.handle, 1, np.array([1000.0]), np.array([ONE_ATM]),
                                               np.ascontiguousarray(Y), 1e-8, 1e-16, 16), "create")
