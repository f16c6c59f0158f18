"""Forward sensitivities of reactor batches on the GPU (sensitivity kernels of
every integrator variant) against the oracle, whose sensitivities are pinned
by finite differences of the solution (test_oracle_sensitivity.py).

Both run the same method (staggered corrector, centred difference-quotient
right-hand sides, sensitivities outside the error test) with the same
Jacobian (analytic species columns: the oracle's analytic mode, as in the
plain integrator's parity tests); they differ in the rounding of the
right-hand sides, which moves the step sequence at the level of the
integration tolerance.  (Against the oracle's difference-quotient Jacobian,
the mode the reference runs, n-dodecane's trace species differ by up to 50
times the state bar below: another valid trajectory at rtol 1e-9, not a
rounding difference.)  Bars, per reactor and parameter:
* the state: as the plain integrator's parity tests (|dT| <= 1e-6 T,
  |dY| <= 1e-9 + 1e-6 |Y|);
* sensitivities, with tight sensitivity tolerances (rtol 1e-7, atol 1e-10):
  1e-4 of the column's largest entry plus 1e-8 absolute, while the solution's
  dependence on p is resolved (the oracle's own bar against finite
  differences; both lie within it of the true dy/dp);
* with the default tolerances (rtol 1e-4, atol 1e-6) the corrector stops
  once a correction is below those weights, so each side carries errors up to
  about atol per step: bar 1e-3 of the column's largest entry plus 10 atol.
The windows end before any thermal runaway: through one, two equally valid
runs (the oracle at rtol 1e-9 and at 1e-9 (1 + 1e-9)) already differ beyond
these bars, by up to 2% in a small column, while before it the same pair stays
below 0.17 of every bar.  The state of a batch with sensitivities equals that
of the same batch without them to 1e-12: the sensitivities never feed back
into the state's steps.
"""
import numpy as np
import pytest

from conftest import mech_file

pytestmark = pytest.mark.gpu
ONE_ATM = 101325.0
H2O2_COMP = (("H2", 0.1), ("OH", 1e-7), ("O2", 0.1), ("AR", 1e-5))
GRI_COMP = (("CO", 1.0), ("H2", 0.1), ("CH4", 0.1), ("H2O", 0.5))


def _Y(sol, comp):
    X = np.zeros(sol.n_species)
    for name, v in comp:
        X[sol.species_index(name)] = v
    return sol.mass_fractions_from_moles(X / X.sum())


TIGHT = (1e-7, 1e-10)


def _compare(oracle_mod, mech, sol, T0, P, Y, cv, params, times, storage=None, stol=TIGHT):
    import cantera_amd as ca
    n = len(T0)
    b = ca.ReactorBatch(sol, T0, P, np.tile(Y, (n, 1)), rtol=1e-9, atol=1e-15, const_volume=cv)
    plain = ca.ReactorBatch(sol, T0, P, np.tile(Y, (n, 1)), rtol=1e-9, atol=1e-15, const_volume=cv)
    if storage is not None:
        b.set_matrix_storage(storage)
        plain.set_matrix_storage(storage)
    for p in params:
        b.add_sensitivity_reaction(p)
    if stol:
        b.set_sensitivity_tolerances(*stol)
        bar, floor = 1e-4, 1e-8
    else:
        bar, floor = 1e-3, 10 * 1e-6
    assert np.all(b.sensitivities() == 0.0)
    om = oracle_mod.Mech(mech_file(mech))
    refs = []
    for s in range(n):
        r = oracle_mod.Reactor(om, T0[s], P, Y, rtol=1e-9, atol=1e-15, const_volume=cv, analytic=True)
        for p in params:
            r.add_sensitivity_reaction(p)
        if stol:
            r.set_sensitivity_tolerances(*stol)
        refs.append(r)
    worst = 0.0
    for t in times:
        print(f"{mech} n={n} params {params} t={t:g}", flush=True)
        b.advance(np.full(n, t))
        plain.advance(np.full(n, t))
        ys, yp = b.state, plain.state
        assert np.allclose(ys, yp, rtol=1e-12, atol=1e-300)
        S = b.sensitivities(normalized=False)
        Sn = b.sensitivities()
        den = np.where(ys == 0.0, 1e-300, ys)
        assert np.allclose(Sn, S / den[:, :, None], rtol=1e-15, atol=0.0)
        for s in range(n):
            refs[s].advance(t)
            yo = refs[s].state
            assert ys[s, 0] == pytest.approx(yo[0], rel=1e-12)
            assert abs(ys[s, 1] - yo[1]) <= 1e-6 * yo[1], (t, s)
            assert np.all(np.abs(ys[s, 2:] - yo[2:]) <= 1e-9 + 1e-6 * np.abs(yo[2:])), (t, s)
            So = refs[s].sensitivities
            assert np.all(S[s, 0] == 0.0)
            for j in range(len(params)):
                sc = np.abs(So[:, j]).max()
                d = float(np.abs(S[s, :, j] - So[:, j]).max())
                if sc > 1e-6:
                    worst = max(worst, d / sc)
                assert d <= bar * sc + floor, (t, s, params[j], d, sc)
    print(f"worst sensitivity difference {worst:.2e}", flush=True)
    return b


def test_h2o2_const_pressure_induction(oracle_mod):
    """Register variant (N = 12): 16 reactors at 900-1000 K, 1 atm, up to
    20 us, before any of them runs away."""
    import cantera_amd as ca
    sol = ca.Solution("h2o2")
    T0 = np.linspace(900.0, 1000.0, 16)
    _compare(oracle_mod, "h2o2", sol, T0, ONE_ATM, _Y(sol, H2O2_COMP), False, [2, 10, 18, 19], [5e-6, 1e-5, 2e-5])


def test_h2o2_const_volume_sensitivities2_setup(oracle_mod):
    """test_sensitivities2's h2o2 reactor (test/python/test_reactor.py:
    2966-3005): 900 K, 1 atm, IdealGasReactor, reactions 2, 10, 18, 19."""
    import cantera_amd as ca
    sol = ca.Solution("h2o2")
    _compare(oracle_mod, "h2o2", sol, np.array([900.0, 905.0, 910.0, 915.0]), ONE_ATM, _Y(sol, H2O2_COMP), True,
             [2, 10, 18, 19], [1e-3, 1e-2, 3e-2])


@pytest.mark.parametrize("storage,stol", [(None, TIGHT), ("global", TIGHT), (None, None)])
def test_gri30_sensitivities1_setup(oracle_mod, storage, stol):
    """test_sensitivities1 (test_reactor.py:2949-2964): gri30, 1300 K, 20 atm,
    reactions 40 and 41, to 0.1 s; LDS-matrix variant and the global-matrix
    one (N = 55); the reference's default sensitivity tolerances too."""
    import cantera_amd as ca
    sol = ca.Solution("gri30")
    b = _compare(oracle_mod, "gri30", sol, np.array([1300.0, 1300.0, 1250.0, 1280.0]), 20 * ONE_ATM,
                 _Y(sol, GRI_COMP), True, [40, 41], [1e-2, 0.1], storage=storage, stol=stol)
    assert b.sensitivities().shape == (4, sol.n_species + 2, 2)


def test_ndodecane_two_slot_variant(oracle_mod):
    """Two state slots per lane (N > 64): n-dodecane, 2 reactors, 3 reactions,
    at 740-750 K, 20 atm, up to 0.8 ms (before the first-stage runaway)."""
    import cantera_amd as ca
    sol = ca.Solution("nDodecane")
    comp = (("c12h26", 1.0), ("o2", 18.5), ("n2", 69.56))
    _compare(oracle_mod, "nDodecane", sol, np.array([740.0, 750.0]), 20 * ONE_ATM, _Y(sol, comp), False,
             [0, 5, 20], [6e-4, 8e-4])


def test_parameter_order_on_gpu():
    """test_parameter_order1 (test_reactor.py:3008-3049) on the GPU."""
    import cantera_amd as ca
    sol = ca.Solution("h2o2")
    Y = _Y(sol, H2O2_COMP)
    out = []
    for order in ([2, 10, 18, 19], [19, 10, 2, 18]):
        b = ca.ReactorBatch(sol, [900.0, 950.0], ONE_ATM, np.tile(Y, (2, 1)), const_volume=True)
        for p in order:
            b.add_sensitivity_reaction(p)
        b.advance(np.full(2, 1e-2))
        out.append(dict(zip(order, np.moveaxis(b.sensitivities(), 2, 0))))
    for p in (2, 10, 18, 19):
        assert out[0][p] == pytest.approx(out[1][p], rel=1e-12, abs=1e-300)


def test_sensitivity_errors():
    import cantera_amd as ca
    sol = ca.Solution("h2o2")
    Y = _Y(sol, H2O2_COMP)
    b = ca.ReactorBatch(sol, [1000.0], ONE_ATM, Y.reshape(1, -1))
    with pytest.raises(ca.CanteraGpuError, match="out of range"):
        b.add_sensitivity_reaction(sol.n_reactions)
    b.advance([1e-6])
    with pytest.raises(ca.CanteraGpuError, match="after the integrator"):
        b.add_sensitivity_reaction(2)
    m = ca.ReactorBatch(sol, [1000.0], ONE_ATM, Y.reshape(1, -1), precision=ca.CG_MIXED)
    with pytest.raises(ca.CanteraGpuError, match="FP64"):
        m.add_sensitivity_reaction(2)
    s = ca.ReactorBatch(sol, [1000.0], ONE_ATM, Y.reshape(1, -1))
    s.add_sensitivity_reaction(2)
    with pytest.raises(ca.CanteraGpuError, match="sensitivity"):
        s.autotune([1e-6])
    with pytest.raises(ca.CanteraGpuError, match="positive"):
        s.set_sensitivity_tolerances(0.0, 1e-6)
