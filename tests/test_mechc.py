"""Mechanism compiler known answers (no GPU needed).

Unit handling is pinned on the reference's own YAML known answers in
test/kinetics/kineticsFromYaml.cpp, compiled here from the same reaction
entries.
"""
import math

import pytest
import yaml

from cantera_amd import mechc

ONE_ATM = 101325.0
GRI_SPECIES = ["H2", "H", "O", "O2", "OH", "H2O", "HO2", "H2O2", "C", "CH", "CH2", "CH2(S)", "CH3", "CH4", "CO",
               "CO2", "HCO", "CH2O", "CH2OH", "CH3O", "CH3OH", "AR", "N2"]
INDEX = {n: i for i, n in enumerate(GRI_SPECIES)}


def _compile(text, file_units=None):
    node = yaml.load(text, Loader=mechc._Yaml12Loader)
    return mechc._compile_reaction(node, INDEX, mechc.UnitSystem.from_node(file_units))


def test_chemically_activated_reaction_local_units():
    """kineticsFromYaml.cpp:457-471: a reaction's own `units` map converts both
    rates; low-P A 282320.078 cm^3/mol/s -> 2.82320078e2 m^3/kmol/s (the
    reaction's order), high-P A 5.88e-14 1/s (one order lower)."""
    r = _compile("{equation: CH3 + OH (+M) <=> CH2O + H2 (+M),"
                 " units: {length: cm, quantity: mol},"
                 " type: chemically-activated,"
                 " high-P-rate-constant: [5.88E-14, 6.721, -3022.227],"
                 " low-P-rate-constant: [282320.078, 1.46878, -3270.56495]}")
    assert r.chem_act and r.rtype == mechc.TYPE_LINDEMANN
    assert r.A == pytest.approx(5.88e-14, rel=1e-15)
    assert r.A0 == pytest.approx(2.82320078e2, rel=1e-15)
    # energies: the reaction-local quantity (mol) gives J/mol -> Ea/R in K
    assert r.EaR == pytest.approx(-3022.227 * 1e3 / mechc.GAS_CONSTANT, rel=1e-14)


def test_reaction_local_units_do_not_leak():
    """The override applies to its own reaction only: the next reaction keeps
    the file defaults (cm, mol, cal/mol as in gri30.yaml)."""
    fu = {"length": "cm", "quantity": "mol", "activation-energy": "cal/mol"}
    a = _compile("{equation: O + H2 <=> H + OH, rate-constant: {A: 3.87e+04, b: 2.7, Ea: 6260.0},"
                 " units: {length: m, quantity: kmol}}", fu)
    b = _compile("{equation: O + H2 <=> H + OH, rate-constant: {A: 3.87e+04, b: 2.7, Ea: 6260.0}}", fu)
    assert a.A == pytest.approx(3.87e4, rel=1e-15)
    assert b.A == pytest.approx(3.87e4 * 1e-3, rel=1e-15)
    # the explicit activation-energy default survives the local override
    assert a.EaR == pytest.approx(b.EaR, rel=1e-15)
    assert b.EaR == pytest.approx(6260.0 * 4184.0 / mechc.GAS_CONSTANT, rel=1e-14)


def test_plog_reaction_local_pressure_units():
    """kineticsFromYaml.cpp:476-500: `units: {pressure: atm}` on the reaction
    makes bare pressures atm; entries with explicit units keep theirs."""
    r = _compile("equation: 'H + CH4 <=> H2 + CH3'\n"
                 "units: {pressure: atm}\n"
                 "type: pressure-dependent-Arrhenius\n"
                 "rate-constants:\n"
                 "- {P: 0.039474, A: 2.720000e+09 cm^3/mol/s, b: 1.2, Ea: 6834.0}\n"
                 "- {P: 1.0 atm, A: 1.260000e+20, b: -1.83, Ea: 15003.0}\n"
                 "- {P: 1.0 atm, A: 1.230000e+04, b: 2.68, Ea: 6335.0}\n"
                 "- {P: 1.01325 MPa, A: 1.680000e+16, b: -0.6, Ea: 14754.0}")
    assert r.rtype == mechc.TYPE_PLOG and len(r.plog) == 4
    assert r.plog[0][0] == pytest.approx(0.039474 * ONE_ATM, abs=1e-6)
    assert r.plog[2][0] == pytest.approx(ONE_ATM, abs=1e-6)
    assert r.plog[3][0] == pytest.approx(10 * ONE_ATM, abs=1e-6)
    assert r.plog[0][1] == pytest.approx(2.72e6, rel=1e-15)
    assert r.plog[3][1] == pytest.approx(1.68e16, rel=1e-15)
    assert r.plog[3][2] == -0.6


def test_energy_default_sets_activation_energy():
    """UnitSystem::setDefaults: without an explicit activation-energy unit,
    Ea is in energy / quantity (kcal/mol here)."""
    r = _compile("{equation: O + H2 <=> H + OH, rate-constant: {A: 1.0, b: 0.0, Ea: 1.5}}",
                 {"energy": "kcal", "quantity": "mol"})
    assert r.EaR == pytest.approx(1.5 * 4184.0 * 1e3 / mechc.GAS_CONSTANT, rel=1e-14)


def test_unknown_units_entry_is_rejected():
    with pytest.raises(mechc.MechanismError):
        _compile("{equation: O + H2 <=> H + OH, rate-constant: {A: 1.0, b: 0.0, Ea: 1.5},"
                 " units: {lenght: cm}}")
    assert math.isfinite(_compile("{equation: O + H2 <=> H + OH, rate-constant: [1.0, 0.0, 0.0]}").A)
