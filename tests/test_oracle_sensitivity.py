"""Pins of the oracle's forward sensitivities (no GPU needed).

The oracle integrates dy/dp for reaction-rate multipliers p (nominal 1) with
the state, as the reference's ReactorNet does through CVODES (staggered
corrector, sensitivities outside the error test, centred difference-quotient
right-hand sides; src/numerics/CVodesIntegrator.cpp:263-279,623-644,
src/zeroD/Reactor.cpp:278-289,384-419).  SUNDIALS is not vendored, so the
restatement is pinned by what the mathematics fixes:

* finite differences of the solution: two more integrations at tolerance
  1e-12 with the multiplier at 1 +- eps give dy/dp to O(eps^2) plus their own
  integration noise (about 1e-12 |y| / eps); the sensitivities agree with them
  to 1e-4 of the column's largest entry on the reference's own test setups
  (test/python/test_reactor.py:2948-3060) while the solution's dependence on
  p is resolved -- before ignition, or with tight sensitivity tolerances
  through it;
* invariants: the mass has zero sensitivity exactly; the species rows sum to
  zero and conserve every element (sum_k Y_k = 1 and the element totals do not
  depend on p), to 1e-6 of the column's scale before ignition;
* parameter order does not change a parameter's column
  (test_reactor.py:3008-3049);
* the sensitivities are zero before the first step (CVodesIntegrator.cpp:
  625-628).

After a thermal runaway the sensitivities keep an O(1e-12 x peak) residue
along the conserved directions that never decays: the error of the
difference quotients through the runaway, uncontrolled when sensitivities are
left out of the error test (the reference's own configuration).  No bar is
placed there.
"""
import numpy as np
import pytest

from conftest import mech_file

ONE_ATM = 101325.0


def _setup(oracle_mod, mech, comp):
    from cantera_amd.workloads import MechInfo
    info = MechInfo(mech_file(mech))
    m = oracle_mod.Mech(mech_file(mech))
    X = np.zeros(m.K)
    for name, v in comp:
        X[info.species_names.index(name)] = v
    X /= X.sum()
    return info, m, info.mass_fractions_from_moles(X)


H2O2_COMP = (("H2", 0.1), ("OH", 1e-7), ("O2", 0.1), ("AR", 1e-5))
GRI_COMP = (("CO", 1.0), ("H2", 0.1), ("CH4", 0.1), ("H2O", 0.5))


def _fd_check(oracle_mod, mech, comp, T0, P, cv, params, fd_params, times, eps, stol=None, bar=1e-4,
              analytic=False):
    info, m, Y = _setup(oracle_mod, mech, comp)
    r = oracle_mod.Reactor(m, T0, P, Y, rtol=1e-9, atol=1e-15, const_volume=cv, analytic=analytic)
    for p in params:
        r.add_sensitivity_reaction(p)
    if stol:
        r.set_sensitivity_tolerances(*stol)
    assert np.all(r.sensitivities == 0.0)
    pert = {}
    for p in fd_params:
        for sg in (1, -1):
            mult = np.ones(m.I)
            mult[p] = 1.0 + sg * eps
            rr = oracle_mod.Reactor(m, T0, P, Y, rtol=1e-12, atol=1e-20, const_volume=cv)
            rr.set_multipliers(mult)
            pert[p, sg] = rr
    A = info.composition / info.molecular_weights[:, None]
    worst = 0.0
    for t in times:
        r.advance(t)
        for rr in pert.values():
            rr.advance(t)
        S = r.sensitivities
        assert S.shape == (m.K + 2, len(params))
        assert np.all(S[0] == 0.0)
        for j, p in enumerate(params):
            if p not in fd_params:
                continue
            fd = (pert[p, 1].state - pert[p, -1].state) / (2 * eps)
            err = float(np.abs(S[:, j] - fd).max() / np.abs(fd).max())
            worst = max(worst, err)
            assert err <= bar, (mech, t, p, err)
            sy = S[2:, j]
            scale = np.abs(sy).max()
            assert abs(sy.sum()) <= 1e-6 * scale, (t, p)
            assert np.abs(sy @ A).max() <= 1e-6 * np.abs(sy[:, None] * A).max(), (t, p)
    return r, worst


def test_h2o2_const_volume_induction(oracle_mod):
    """test_sensitivities2's h2o2 reactor (900 K, 1 atm, IdealGasReactor,
    reaction 18 among its parameters), through 50 ms of slow induction."""
    _, worst = _fd_check(oracle_mod, "h2o2", H2O2_COMP, 900.0, ONE_ATM, True, [2, 10, 18, 19], [10, 18],
                         [1e-3, 1e-2, 5e-2], 1e-4)
    print(f"h2o2 900 K: worst {worst:.2e}")


def test_analytic_jacobian_mode(oracle_mod):
    """The same pins with the analytic-species-column Jacobian (the mode the
    GPU integrator runs): the sensitivities use the state's iteration matrix
    whichever Jacobian formed it.  gri30 with the sensitivity tolerances of
    the GPU parity tests (1e-7, 1e-10): at the default atol of 1e-6 a column
    whose entries are ~1e-4 is converged only to ~2e-4 of its size here."""
    _fd_check(oracle_mod, "h2o2", H2O2_COMP, 1000.0, ONE_ATM, False, [2, 10, 18], [2, 10, 18], [2e-5, 5e-5], 1e-4,
              analytic=True)
    _fd_check(oracle_mod, "gri30", GRI_COMP, 1300.0, 20 * ONE_ATM, True, [40], [40], [1e-2, 0.1], 1e-3,
              stol=(1e-7, 1e-10), analytic=True)


def test_h2o2_const_pressure_to_ignition(oracle_mod):
    """1000 K constant pressure up to the start of the runaway (T 1007 K at
    50 us), and with tight sensitivity tolerances across it (T 3035 K at
    100 us, 1e-3 bar: the residue described above)."""
    _fd_check(oracle_mod, "h2o2", H2O2_COMP, 1000.0, ONE_ATM, False, [2, 10, 18, 19], [2, 10, 18], [2e-5, 5e-5], 1e-4)
    info, m, Y = _setup(oracle_mod, "h2o2", H2O2_COMP)
    r = oracle_mod.Reactor(m, 1000.0, ONE_ATM, Y, rtol=1e-9, atol=1e-15)
    r.add_sensitivity_reaction(10)
    r.set_sensitivity_tolerances(1e-8, 1e-10)
    eps = 1e-4
    ys = []
    for sg in (1, -1):
        mult = np.ones(m.I)
        mult[10] = 1.0 + sg * eps
        rr = oracle_mod.Reactor(m, 1000.0, ONE_ATM, Y, rtol=1e-12, atol=1e-20)
        rr.set_multipliers(mult)
        rr.advance(1e-4)
        ys.append(rr.state)
    r.advance(1e-4)
    assert r.state[1] > 3000.0
    fd = (ys[0] - ys[1]) / (2 * eps)
    assert np.abs(r.sensitivities[:, 0] - fd).max() <= 1e-3 * np.abs(fd).max()


def test_gri30_sensitivities1_setup(oracle_mod):
    """test_sensitivities1 (test_reactor.py:2949-2964): gri30, 1300 K, 20 atm,
    CO:1 H2:0.1 CH4:0.1 H2O:0.5, IdealGasReactor, reactions 40 and 41, advance
    to 0.1 s; shape (n_vars, n_params) and reaction 40's column against finite
    differences (eps 1e-3; reaction 41's column, ~1e-7, is below the noise of
    the difference over T)."""
    r, worst = _fd_check(oracle_mod, "gri30", GRI_COMP, 1300.0, 20 * ONE_ATM, True, [40, 41], [40], [1e-2, 0.1],
                         1e-3)
    assert r.sensitivities.shape == (53 + 2, 2)
    print(f"gri30 1300 K: worst {worst:.2e}")


def test_parameter_order(oracle_mod):
    """test_parameter_order1 (test_reactor.py:3008-3049): the same reactions
    added in another order give the same columns, permuted."""
    _, m, Y = _setup(oracle_mod, "h2o2", H2O2_COMP)
    out = []
    for order in ([2, 10, 18, 19], [19, 10, 2, 18]):
        r = oracle_mod.Reactor(m, 900.0, ONE_ATM, Y, rtol=1e-9, atol=1e-15, const_volume=True)
        for p in order:
            r.add_sensitivity_reaction(p)
        while r.state[1] < 910.0:
            r.step()
        out.append(dict(zip(order, r.sensitivities.T)))
    for p in (2, 10, 18, 19):
        assert out[0][p] == pytest.approx(out[1][p], rel=1e-12, abs=1e-300)


def test_add_after_start_rejected(oracle_mod):
    _, m, Y = _setup(oracle_mod, "h2o2", H2O2_COMP)
    r = oracle_mod.Reactor(m, 1000.0, ONE_ATM, Y)
    r.advance(1e-6)
    with pytest.raises(RuntimeError):
        r.add_sensitivity_reaction(2)
    r2 = oracle_mod.Reactor(m, 1000.0, ONE_ATM, Y)
    with pytest.raises(RuntimeError):
        r2.add_sensitivity_reaction(m.I)
