"""Chemkin-format input (cantera_amd.ckc) against the reference's own
conversion tests (test/python/test_convert.py, Testck2yaml): each Chemkin
fixture of test/data is compiled and compared with its Cantera-YAML
counterpart compiled by mechc -- elements, species, compositions and
stoichiometric coefficients equal; forward and reverse rate constants at the
reference's temperatures and pressures within its bar (1e-8 relative; 1e-7
for the units test); standard-state cp, h, s within 1e-7.  Rate constants
and thermo come from the oracle loaded with each compiled table, so the
check covers the compiled numbers, not the parser's intermediate text.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, mech_file

CK = os.path.join(GOLDEN, "chemkin")


def _compile_ck(tmp_path, inp, thermo=None):
    from cantera_amd import ckc, mechc
    m = ckc.compile_chemkin(os.path.join(CK, inp), os.path.join(CK, thermo) if thermo else None)
    out = tmp_path / (inp + ".mech")
    mechc.write_mech(m, str(out))
    return str(out), m


def _compile_yaml(tmp_path, yml):
    from cantera_amd import mechc
    m = mechc.compile_mechanism(os.path.join(CK, yml))
    out = tmp_path / (yml + ".mech")
    mechc.write_mech(m, str(out))
    return str(out), m


def _check_structure(a, b):
    assert a["elements"] == b["elements"]
    assert [s.name for s in a["species"]] == [s.name for s in b["species"]]
    assert [s.composition for s in a["species"]] == [s.composition for s in b["species"]]
    assert len(a["reactions"]) == len(b["reactions"])
    for ra, rb in zip(a["reactions"], b["reactions"]):
        assert ra.reac == rb.reac and ra.prod == rb.prod, ra.equation
        assert ra.reversible == rb.reversible and ra.rtype == rb.rtype, ra.equation


def _check_kinetics(oracle_mod, pa, pb, temps, pressures, tol=1e-8):
    ma, mb = oracle_mod.Mech(pa), oracle_mod.Mech(pb)
    Y = np.random.default_rng(3).random(ma.K) + 0.1
    Y /= Y.sum()
    n = 0
    for T in temps:
        for P in pressures:
            ra, rb = ma.rates(T, P, Y), mb.rates(T, P, Y)
            kfa, kfb = ra["kf"], rb["kf"]
            kra, krb = kfa * ra["rkcn"], kfb * rb["rkcn"]
            assert np.allclose(kfa, kfb, rtol=tol, atol=0), (T, P)
            assert np.allclose(kra, krb, rtol=tol, atol=0), (T, P)
            n += len(kfa)
    return ma, mb, n


def _check_thermo(ma, mb, temps):
    for T in temps:
        for xa, xb in zip(ma.thermo(T), mb.thermo(T)):
            assert np.allclose(xa, xb, rtol=1e-7, atol=0), T


def test_gri30(tmp_path, oracle_mod):
    """test_convert.py:89-94: gri30.inp + gri30_thermo.dat against gri30.yaml
    (here its compiled tables), k_f and k_r at 300 and 1500 K, 5e3 - 2e6 Pa."""
    from cantera_amd.workloads import MechInfo
    pa, a = _compile_ck(tmp_path, "gri30.inp", "gri30_thermo.dat")
    b = MechInfo(mech_file("gri30"))
    assert a["elements"] == b.elements and [s.name for s in a["species"]] == b.species_names
    assert np.array_equal(MechInfo(pa).composition, b.composition)
    ma, mb, n = _check_kinetics(oracle_mod, pa, mech_file("gri30"), [300, 1500], [5e3, 1e5, 2e6])
    _check_thermo(ma, mb, [300, 1100, 2500])
    assert np.array_equal(ma.mw, mb.mw)
    print(f"gri30: {n} rate constants agree", flush=True)


@pytest.mark.parametrize("inp,thermo,ref,temps,pressures", [
    ("pdep-test.inp", None, "pdep-test", [300, 800, 1450, 2800], [5e3, 1e5, 2e6]),
    ("sri-falloff.inp", "dummy-thermo.dat", "sri-falloff", [300, 800, 1450, 2800], [5e3, 1e5, 2e6]),
    ("chemically-activated-reaction.inp", None, "chemically-activated-reaction", [300, 800, 1450, 2800],
     [5e3, 1e5, 2e6, 1e7]),
])
def test_against_committed_tables(tmp_path, oracle_mod, inp, thermo, ref, temps, pressures):
    """test_convert.py:102-105, 291-300: PLOG / Chebyshev (local units, M
    dropped, duplicate pressures, negative A), SRI, chemically activated."""
    pa, _ = _compile_ck(tmp_path, inp, thermo)
    _check_kinetics(oracle_mod, pa, mech_file(ref), temps, pressures)


@pytest.mark.parametrize("inp,yml", [
    ("explicit-third-bodies.inp", "explicit-third-bodies.yaml"),
    ("explicit-reverse-rate.inp", "explicit-reverse-rate.yaml"),
    ("explicit-forward-order.inp", "explicit-forward-order.yaml"),
    ("negative-order.inp", "negative-order.yaml"),
    ("nonreactant-order.inp", "nonreactant-order.yaml"),
])
def test_against_yaml_fixtures(tmp_path, oracle_mod, inp, yml):
    """test_convert.py:307-354: explicit colliders and efficiencies (including
    a species named SP)X), explicit reverse rates, FORD orders (negative and
    non-reactant ones flagged)."""
    pa, a = _compile_ck(tmp_path, inp, "dummy-thermo.dat")
    pb, b = _compile_yaml(tmp_path, yml)
    _check_structure(a, b)
    _check_kinetics(oracle_mod, pa, pb, [300, 800, 1450, 2800], [5e3, 1e5, 2e6])


def test_explicit_reverse_rate_split(tmp_path, oracle_mod):
    """test_convert.py:312-333: explicit reverse rates become pairs of
    irreversible reactions with swapped sides; a zero reverse A keeps only the
    forward one; five reactions in all, none with a reverse rate."""
    pa, a = _compile_ck(tmp_path, "explicit-reverse-rate.inp", "dummy-thermo.dat")
    r = a["reactions"]
    assert len(r) == 5 and not any(x.reversible for x in r)
    assert [k for k, _, _ in r[0].reac] == [k for k, _ in r[1].prod]
    assert [k for k, _ in r[0].prod] == [k for k, _, _ in r[1].reac]
    m = oracle_mod.Mech(pa)
    assert np.all(m.rates(1000.0, 1e5, np.full(m.K, 1.0 / m.K))["rkcn"] == 0.0)


def test_reaction_units(tmp_path, oracle_mod):
    """test_convert.py:386-391: the same reactions written with the default
    units and with KELVIN / MOLECULES plus per-reaction UNITS give the same
    rate constants (1e-7)."""
    pa, a = _compile_ck(tmp_path, "units-default.inp", "dummy-thermo.dat")
    pb, b = _compile_ck(tmp_path, "units-custom.inp", "dummy-thermo.dat")
    _check_structure(a, b)
    _check_kinetics(oracle_mod, pa, pb, [300, 800, 1450, 2800], [5e0, 5e3, 1e5, 2e6, 1e8], tol=1e-7)


def test_float_stoich_coeffs(tmp_path):
    """test_convert.py:393-402: fractional coefficients written before names."""
    _, a = _compile_ck(tmp_path, "float-stoich.inp", "dummy-thermo.dat")
    names = [s.name for s in a["species"]]
    r0 = a["reactions"][0]
    reac = {names[k]: s for k, s, _ in r0.reac}
    prod = {names[k]: s for k, s in r0.prod}
    assert reac == {"R1A": 1.5, "R1B": 0.5} and prod == {"H": 1.0, "P1": 1.0}
    r1 = a["reactions"][1]
    assert {names[k]: s for k, s in r1.prod} == {"R1A": 0.33, "R1B": 1.67}


def test_errors(tmp_path):
    from cantera_amd import ckc
    from cantera_amd.mechc import MechanismError
    bad = tmp_path / "bad.inp"
    bad.write_text("ELEM H C END\nSPECIES H R1A R1B P1 END\nREACTIONS\nR1A+R1B(+M)=P1(+M) 1e12 0 0\nEND\n")
    with pytest.raises(MechanismError, match="pressure dependence"):
        ckc.read_chemkin(str(bad), os.path.join(CK, "dummy-thermo.dat"))
    bad.write_text("ELEM H END\nSPECIES H END\nREACTIONS\nH+X=H 1e12 0 0\nEND\n")
    with pytest.raises(MechanismError, match="species"):
        ckc.read_chemkin(str(bad), os.path.join(CK, "dummy-thermo.dat"))
    bad.write_text("ELEM H END\nSPECIES H END\nREACTIONS\nH=H 1e12 0 0\n  XSMI/1/\nEND\n")
    with pytest.raises(MechanismError, match="unsupported"):
        ckc.read_chemkin(str(bad), os.path.join(CK, "dummy-thermo.dat"))
