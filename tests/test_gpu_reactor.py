"""Batched constant-pressure reactor integration on the GPU.

Parity bars (stated per test):
* kin1 (h2o2, T0 = 1001 K, 1 atm, H2:2 O2:1 N2:4, 100 outputs every 1e-5 s):
  the reference's own comparator on test_problems/cxx_samples/kin1_blessed.csv,
  |a - b| / (max(|a|, |b|) + 1e-14) < 2e-4 on 6-significant-figure output
  (test_problems/SConscript:236-237, site_scons/buildutils.py:1016-1071).
* GPU against the oracle running the same algorithm (analytic Jacobian):
  the two differ only by floating-point rounding, so trajectories agree far
  below the integrator tolerance: |dT| <= 1e-6 T and |dY| <= 1e-9 + 1e-6 |Y|.
* GPU against the oracle in its CVODES-like mode (difference-quotient
  Jacobian): within the integrator tolerance, |dT| <= 1e-4 T and
  |dY| <= 1e-7 + 1e-3 |Y|.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, mech_file

pytestmark = pytest.mark.gpu


def kin1_inputs(info):
    X = np.zeros(info.n_species)
    X[info.species_index("H2")] = 2.0
    X[info.species_index("O2")] = 1.0
    X[info.species_index("N2")] = 4.0
    X /= X.sum()
    return info.mass_fractions_from_moles(X)


def csv_rows(info, t, y):
    """Rows as kinetics1 writes them: t, T, rho, P, X_k (6 significant figures)."""
    Y = y[2:]
    mmw = 1.0 / np.sum(Y / info.molecular_weights)
    X = Y / info.molecular_weights * mmw
    rho = 101325.0 * mmw / (6.02214076e26 * 1.380649e-23 * y[1])
    return [float("%.6g" % v) for v in [t, y[1], rho, 101325.0, *X]]


def test_kin1_blessed_trajectory():
    import cantera_amd as ca
    from cantera_amd.workloads import MechInfo
    info = MechInfo(mech_file("h2o2"))
    sol = ca.Solution("h2o2")
    Y0 = kin1_inputs(info)
    rb = ca.ReactorBatch(sol, [1001.0], 101325.0, Y0.reshape(1, -1))
    rows = [csv_rows(info, 0.0, rb.state[0])]
    for i in range(1, 101):
        rb.advance(i * 1e-5)
        rows.append(csv_rows(info, i * 1e-5, rb.state[0]))
    ours = np.array(rows)
    ref = np.genfromtxt(os.path.join(GOLDEN, "kin1_blessed.csv"), delimiter=",", skip_header=1)
    rel = np.abs(ours - ref) / (np.maximum(np.abs(ours), np.abs(ref)) + 1e-14)
    print("kin1 max comparator error", np.nanmax(rel), "stats", {k: int(v[0]) for k, v in rb.stats.items()})
    assert np.nanmax(rel) < 2e-4


def _compare(mech, wl, nout, analytic, tolT, tolY_abs, tolY_rel, oracle_mod, precision=0, const_volume=False):
    import cantera_amd as ca
    sol = ca.Solution(mech)
    om = oracle_mod.Mech(mech_file(mech))
    n = len(wl.T0)
    rb = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol, precision=precision,
                         const_volume=const_volume)
    gpu = np.zeros((n, nout + 1, sol.n_species + 2))
    gpu[:, 0] = rb.state
    for k in range(1, nout + 1):
        rb.advance(k * wl.dt)
        gpu[:, k] = rb.state
    ref, stats = om.run_reactors(wl.T0, wl.P, wl.Y0, wl.dt, nout, wl.rtol, wl.atol, analytic=analytic, nthreads=8,
                                 const_volume=const_volume)
    dT = np.abs(gpu[:, :, 1] - ref[:, :, 1]) / ref[:, :, 1]
    dY = np.abs(gpu[:, :, 2:] - ref[:, :, 2:])
    okY = dY <= tolY_abs + tolY_rel * np.abs(ref[:, :, 2:])
    gs = rb.stats
    print(f"{mech} analytic={analytic} prec={precision}: max dT/T {dT.max():.3e}, max dY {dY.max():.3e}, "
          f"gpu steps {gs['steps'].mean():.1f} oracle steps {stats[:, 0].mean():.1f}", flush=True)
    assert dT.max() <= tolT
    assert okY.all()
    return gpu, ref


def test_h2o2_batch_matches_oracle_same_algorithm(oracle_mod):
    from cantera_amd.workloads import MechInfo, config1
    wl = config1(MechInfo(mech_file("h2o2")), n=96)
    _compare("h2o2", wl, 100, True, 1e-6, 1e-9, 1e-6, oracle_mod)


def test_h2o2_batch_within_tolerance_of_cvodes_mode(oracle_mod):
    from cantera_amd.workloads import MechInfo, config1
    wl = config1(MechInfo(mech_file("h2o2")), n=96)
    _compare("h2o2", wl, 100, False, 1e-4, 1e-7, 1e-3, oracle_mod)


def test_gri30_batch_matches_oracle(oracle_mod):
    from cantera_amd.workloads import MechInfo, config3
    wl = config3(MechInfo(mech_file("gri30")), n=48, nsteps=10)
    _compare("gri30", wl, 10, True, 1e-6, 1e-9, 1e-6, oracle_mod)


def test_ndodecane_batch_matches_oracle(oracle_mod):
    """N = 102: two state slots per lane, iteration matrix in LDS."""
    from cantera_amd.workloads import MechInfo, config5
    wl = config5(MechInfo(mech_file("nDodecane")), n=8, nsteps=5)
    _compare("nDodecane", wl, 5, True, 1e-6, 1e-9, 1e-6, oracle_mod)


def test_pdep_batch_matches_oracle(oracle_mod):
    """PLOG and Chebyshev rate constants inside the integrator and its
    Jacobian (pdep-test.yaml), pressures from 0.01 to 90 atm across the PLOG
    nodes and the Chebyshev range.  The reactants are 1% each in P1, P4 and
    P2A: undiluted, this artificial mechanism runs away within nanoseconds and
    trajectories then amplify rounding differences beyond the bar."""
    from cantera_amd.workloads import MechInfo, ReactorWorkload
    info = MechInfo(mech_file("pdep-test"))
    g = np.random.default_rng(11)
    n = 40
    X = np.zeros((n, info.n_species))
    for name in ("H", "R1A", "R1B", "R2", "R3", "R4", "R5", "R6"):
        X[:, info.species_index(name)] = 0.01 * g.uniform(0.5, 1.5, n)
    for name in ("P1", "P4", "P2A"):
        X[:, info.species_index(name)] += g.uniform(0.5, 1.5, n)
    X /= X.sum(axis=1, keepdims=True)
    P = 101325.0 * np.exp(g.uniform(np.log(0.01), np.log(90.0), n))
    wl = ReactorWorkload("pdep-test IGCPR", g.uniform(800.0, 1200.0, n), P,
                         info.mass_fractions_from_moles(X), np.full(n, 1e-6), 10, 1e-8, 1e-15)
    _compare("pdep-test", wl, 10, True, 1e-6, 1e-9, 1e-6, oracle_mod)


def test_const_volume_batch_matches_oracle(oracle_mod):
    """IdealGasReactor (closed, rigid, V = 1 m^3) batches against the oracle's
    constant-volume reactor running the same algorithm; the density (mass per
    unit volume) stays at its initial value and the specific internal energy is
    conserved along the GPU trajectories."""
    from cantera_amd.workloads import MechInfo, config3
    info = MechInfo(mech_file("gri30"))
    wl = config3(info, n=48, nsteps=10)
    gpu, ref = _compare("gri30", wl, 10, True, 1e-6, 1e-9, 1e-6, oracle_mod, const_volume=True)
    assert np.all(gpu[:, :, 0] == gpu[:, :1, 0])
    om = oracle_mod.Mech(mech_file("gri30"))
    R = 6.02214076e26 * 1.380649e-23

    def u_mass(y):
        _, h, _ = om.thermo(y[1])
        return float(np.sum(y[2:] / om.mw * (h - 1.0) * R * y[1]))

    for r in range(0, 48, 7):
        u0 = u_mass(gpu[r, 0])
        for k in range(1, 11):
            assert abs(u_mass(gpu[r, k]) - u0) <= 1e-6 * abs(u0) + 10.0, (r, k)


def test_gri30_mixed_precision_within_tolerance(oracle_mod):
    """FP32 iteration matrix (FP64 state, residual and error test): the Newton
    corrections converge to the same FP64 solution, so trajectories stay within
    the integrator tolerance bar of the CVODES-mode comparison."""
    from cantera_amd.workloads import MechInfo, config3
    wl = config3(MechInfo(mech_file("gri30")), n=48, nsteps=10)
    _compare("gri30", wl, 10, True, 1e-4, 1e-7, 1e-3, oracle_mod, precision=1)


def test_ndodecane_mixed_precision_within_tolerance(oracle_mod):
    """N = 102 with the FP32 iteration matrix in LDS (two state slots per
    lane): within the CVODES-mode bar of the oracle, as for gri30."""
    from cantera_amd.workloads import MechInfo, config5
    wl = config5(MechInfo(mech_file("nDodecane")), n=8, nsteps=5)
    _compare("nDodecane", wl, 5, True, 1e-4, 1e-7, 1e-3, oracle_mod, precision=1)


def test_schedule_does_not_change_results():
    """Longest-first and index-order work queues integrate the same reactors;
    only the Jacobian's atomic summation order can differ between runs."""
    import cantera_amd as ca
    from cantera_amd.workloads import MechInfo, config3
    wl = config3(MechInfo(mech_file("gri30")), n=96, nsteps=6)
    sol = ca.Solution("gri30")
    out = []
    for lpt in (True, False):
        rb = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol)
        rb.set_schedule(lpt)
        for k in range(1, 7):
            rb.advance(k * wl.dt)
        assert np.all(rb.status == 0)
        out.append(rb.state)
    dT = np.abs(out[0][:, 1] - out[1][:, 1]) / out[1][:, 1]
    assert dT.max() <= 1e-6


def test_matrix_storage_does_not_change_results():
    """Iteration matrix in LDS or in the reactor's global slot: same arithmetic."""
    import cantera_amd as ca
    from cantera_amd.workloads import MechInfo, config3
    wl = config3(MechInfo(mech_file("gri30")), n=64, nsteps=6)
    sol = ca.Solution("gri30")
    out = []
    for storage in ("lds", "global"):
        rb = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol)
        rb.set_matrix_storage(storage)
        for k in range(1, 7):
            rb.advance(k * wl.dt)
        assert np.all(rb.status == 0)
        out.append(rb.state)
    assert np.max(np.abs(out[0][:, 1] - out[1][:, 1]) / out[1][:, 1]) <= 1e-6


def test_graph_replay_matches_eager_advances():
    """cg_reactors_advance_graph captures 4 integrator launches once and replays
    them; three replays with new output times each, with eager work (another
    batch's advances) and synchronisation between them, each agree with the
    same advances done eagerly.  Bar as the schedule test (only atomic
    summation order can differ)."""
    import cantera_amd as ca
    from cantera_amd.workloads import MechInfo, config3
    wl = config3(MechInfo(mech_file("gri30")), n=96, nsteps=12)
    sol = ca.Solution("gri30")
    eager = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol)
    graph = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol)
    K = 4
    for r in range(3):
        times = np.stack([(r * K + k + 1) * wl.dt for k in range(K)])
        for k in range(K):
            eager.advance(times[k])
        graph.advance_steps(times)
        ya, yb = eager.state, graph.state
        assert np.all(graph.status == 0)
        dT = np.abs(ya[:, 1] - yb[:, 1]) / ya[:, 1]
        dY = np.abs(ya[:, 2:] - yb[:, 2:]).max()
        print(f"replay {r}: max rel dT {dT.max():.2e}, max |dY| {dY:.2e}", flush=True)
        assert dT.max() <= 1e-6
        assert dY <= 1e-7


def test_autotune_restores_state_and_keeps_results():
    """cg_reactors_autotune runs every candidate launch from the same state and
    puts that state back: the state right after tuning is bit-identical to the
    one before, and later advances agree with an untuned batch (bar as the
    schedule test)."""
    import cantera_amd as ca
    from cantera_amd.workloads import MechInfo, config3
    wl = config3(MechInfo(mech_file("gri30")), n=128, nsteps=6)
    sol = ca.Solution("gri30")
    tuned = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol)
    plain = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol)
    tuned.advance(wl.dt)
    plain.advance(wl.dt)
    before = tuned.state
    choice, times = tuned.autotune(2 * wl.dt)
    print("autotune choice", choice, "times [ms]", times, flush=True)
    assert choice[0] in (1, 2) and choice[1] in (0, 4)
    assert np.any(times > 0)
    assert np.array_equal(tuned.state, before)
    for k in range(2, 7):
        tuned.advance(k * wl.dt)
        plain.advance(k * wl.dt)
    assert np.all(tuned.status == 0)
    ya, yb = plain.state, tuned.state
    assert (np.abs(ya[:, 1] - yb[:, 1]) / ya[:, 1]).max() <= 1e-6


def test_reactor_invariants():
    """Mass fractions keep summing to 1 and enthalpy is conserved (adiabatic,
    isobaric): checked through the elements and the state after 50 outputs."""
    import cantera_amd as ca
    from cantera_amd.workloads import MechInfo, config3
    info = MechInfo(mech_file("gri30"))
    wl = config3(info, n=64, nsteps=50)
    sol = ca.Solution("gri30")
    rb = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol)
    E0 = wl.Y0 / info.molecular_weights @ info.composition
    for k in range(1, 51):
        rb.advance(k * wl.dt)
    y = rb.state
    assert np.all(rb.status == 0)
    assert np.max(np.abs(y[:, 2:].sum(axis=1) - 1.0)) < 1e-7
    E1 = y[:, 2:] / info.molecular_weights @ info.composition
    assert np.max(np.abs(E1 - E0) / (np.abs(E0) + 1e-12)) < 1e-6


def test_gri30_config3_256_reactors_all_steps(oracle_mod):
    """256 reactors of the headline workload (T0 900-2000 K, phi 0.5-3, 1-40
    atm, per-reactor dt 1e-6..1e-4 s) over all 100 output steps, including its
    stiff tail, against the oracle running the same algorithm.  Over 100 steps
    through ignition a last-bit difference can flip one step-size or
    Jacobian-reuse decision; from then on the two runs differ by the
    integrator's local error, amplified by the ignition's sensitivity, so the
    bar is the integrator-tolerance one (|dT| <= 1e-4 T, |dY| <= 1e-7 + 1e-3
    |Y|).  At least 90% of the reactors must still meet the same-algorithm bar
    (|dT| <= 1e-6 T, |dY| <= 1e-9 + 1e-6 |Y|) at every step."""
    from cantera_amd.workloads import MechInfo, config3
    wl = config3(MechInfo(mech_file("gri30")), n=256, nsteps=100)
    gpu, ref = _compare("gri30", wl, 100, True, 1e-4, 1e-7, 1e-3, oracle_mod)
    tight = ((np.abs(gpu[:, :, 1] - ref[:, :, 1]) <= 1e-6 * ref[:, :, 1]).all(axis=1)
             & (np.abs(gpu[:, :, 2:] - ref[:, :, 2:]) <= 1e-9 + 1e-6 * np.abs(ref[:, :, 2:])).all(axis=(1, 2)))
    print(f"{tight.sum()} of 256 reactors within the same-algorithm bar at all 100 steps", flush=True)
    assert tight.mean() >= 0.9


def test_const_volume_mixed_and_cvodes_mode(oracle_mod):
    """Constant-volume batches in the mixed form (FP32 iteration matrix) and
    against the oracle's CVODES-mode (difference-quotient Jacobian) run: both
    within the integrator-tolerance bar used for constant pressure."""
    from cantera_amd.workloads import MechInfo, config3
    wl = config3(MechInfo(mech_file("gri30")), n=48, nsteps=10)
    _compare("gri30", wl, 10, True, 1e-4, 1e-7, 1e-3, oracle_mod, precision=1, const_volume=True)
    _compare("gri30", wl, 10, False, 1e-4, 1e-7, 1e-3, oracle_mod, const_volume=True)


def test_max_steps_failure_status():
    """A reactor that needs more than max_steps internal steps to reach its
    output time stops with status -22 (CVODES' too-much-work flag, which the
    reference reports as a CanteraError, CVodesIntegrator.cpp:491-512); the
    call returns an error, the failed reactors are left untouched by later
    advances, and other batches keep working."""
    import cantera_amd as ca
    from cantera_amd.workloads import MechInfo, config3
    wl = config3(MechInfo(mech_file("gri30")), n=16, nsteps=2)
    sol = ca.Solution("gri30")
    rb = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol)
    rb.set_options(max_steps=3)
    with pytest.raises(ca.CanteraGpuError):
        rb.advance(1e-3)
    st = rb.status
    print("status after a 3-step limit:", st.tolist(), flush=True)
    assert np.all(st == -22)
    ctl_t = rb.state
    with pytest.raises(ca.CanteraGpuError):
        rb.advance(2e-3)
    assert np.array_equal(rb.state, ctl_t)
    ok = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol)
    ok.advance(1e-5)
    assert np.all(ok.status == 0)


def test_invalid_initial_state_is_rejected():
    """Phase::setTemperature / setDensity reject non-positive values; the
    batch constructor also rejects non-finite inputs, before any kernel runs."""
    import cantera_amd as ca
    from cantera_amd.workloads import MechInfo, config3
    wl = config3(MechInfo(mech_file("gri30")), n=4, nsteps=1)
    sol = ca.Solution("gri30")
    for T0, P in ((-300.0, 101325.0), (np.nan, 101325.0), (1000.0, 0.0)):
        T = wl.T0.copy()
        Pa = wl.P.copy()
        T[2], Pa[2] = T0, P
        with pytest.raises(ca.CanteraGpuError):
            ca.ReactorBatch(sol, T, Pa, wl.Y0)


@pytest.mark.parametrize("const_volume", [False, True])
def test_rates_at_reactor_states(const_volume, oracle_mod):
    """cg_reactors_net_production_rates_device: the rates at each reactor's
    current state (kinetics1's getNetProductionRates after every advance)
    agree with the oracle at those states (bar as the batched rates test);
    for constant volume the state is (T, rho) with P from the ideal gas law."""
    import cantera_amd as ca
    from cantera_amd.workloads import MechInfo, config1
    from conftest import parse_stoich
    info = MechInfo(mech_file("h2o2"))
    sol = ca.Solution("h2o2")
    om = oracle_mod.Mech(mech_file("h2o2"))
    wl = config1(info, n=200)
    rb = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol, const_volume=const_volume)
    nu_r, nu_p = parse_stoich(mech_file("h2o2"))
    absnu = nu_r + nu_p
    R = 6.02214076e26 * 1.380649e-23
    for k in (1, 7, 40):
        rb.advance(k * wl.dt)
        y = rb.state
        w = rb.net_production_rates
        worst = 0.0
        for s in range(0, 200, 9):
            P = wl.P[s]
            if const_volume:
                Wbar = 1.0 / np.sum(y[s, 2:] / info.molecular_weights)
                P = y[s, 0] / 1.0 * R * y[s, 1] / Wbar      # rho = m / V, V = 1 m^3
            r = om.rates(y[s, 1], P, y[s, 2:])
            scale = absnu @ (np.abs(r["ropf"]) + np.abs(r["ropr"])) + 1e-300
            worst = max(worst, float((np.abs(w[:, s] - r["wdot"]) / scale).max()))
        print(f"const_volume={const_volume} output {k}: max normalised error {worst:.2e}", flush=True)
        assert worst <= 1e-10
