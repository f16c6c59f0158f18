"""Mechanism compiler: Cantera-schema YAML (ideal gas, NASA7, bulk kinetics) -> flat tables.

This is the one-time host step that replaces Cantera's mechanism assembly
(``newSolution`` -> ``BulkKinetics::addReaction``; reference files
``src/base/Solution.cpp:242-378``, ``src/kinetics/BulkKinetics.cpp:16-84``,
``src/kinetics/Kinetics.cpp:709-800``).  It reads the YAML with PyYAML's
SafeLoader and writes a line-oriented ``.mech`` text file that the C++ host
library (``cg_mech_load``) turns into device tables.

What is restated from the reference, and where:

* equation parsing and implicit third bodies -- ``Reaction.cpp:389-555`` and
  ``parseReactionEquation`` ``Reaction.cpp:889-955``.  Species maps are ordered
  by name (the reference stores reactants/products in ``std::map<string,double>``),
  so the reactant order written here is the reference's ``rk`` order.
* explicit colliders (``H + O2 + AR <=> HO2 + AR``) become three-body reactions
  with default efficiency 0 and ``{AR: 1}`` -- ``ThirdBody::setName``
  ``Reaction.cpp:775-801``.
* falloff third bodies are not mass-action -- ``Reaction.cpp:173-181``.
* rate-constant units -- ``Reaction::calculateRateCoeffUnits``
  ``Reaction.cpp:576-614``; ``UnitSystem::convertTo`` ``Units.cpp:562-571``;
  activation energies ``UnitSystem::convertActivationEnergyTo``
  ``Units.cpp:713-726`` (``Ea/R`` in K with R = Avogadro*Boltzmann,
  ``ct_defs.h:84-123``).
* molecular weights -- ``Phase::addSpecies`` ``Phase.cpp:835-839`` with the
  atomic-weight table of ``Elements.cpp:65-82``; summed in phase element order.
* NASA7 two-range data -- ``setupNasaPoly`` ``SpeciesThermoFactory.cpp:88-106``.
* Troe coefficients -- ``TroeRate::setFalloffCoeffs`` ``Falloff.cpp:235-270``
  (1/T3 and 1/T1 become +inf when |T| < 1e-300; T2 = 0 means "no T2 term").
* third-body efficiencies are stored as (eff - default), sorted by species
  index -- ``ThirdBodyCalc::install`` ``ThirdBodyCalc.h:23-44``.
* delta-n -- ``BulkKinetics::addReaction`` ``BulkKinetics.cpp:27-35``.
* pressure-dependent Arrhenius (PLOG) -- ``PlogRate::setRates``
  ``PlogRate.cpp:92-123``: entries sorted by pressure (equal pressures keep
  their input order, as in the reference's multimap), each A converted with
  the reaction's rate units; negative A allowed inside a group.
* Chebyshev -- ``ChebyshevRate::setParameters`` ``ChebyshevRate.cpp:62-94``:
  temperature and pressure ranges converted to K and Pa, and log10 of the
  rate-unit conversion of 1.0 added to coefficient (0, 0); reaction-local
  ``units`` are honoured.
* SRI -- ``SriRate::setFalloffCoeffs`` ``Falloff.cpp:347-375``: 3 or 5
  coefficients (d = 1, e = 0 when absent), C and D must not be negative.
* Tsang -- ``TsangRate::setFalloffCoeffs`` ``Falloff.cpp:454-469``: A, and
  B = 0 when absent.
* chemically activated falloff (``type: chemically-activated``) -- the falloff
  record plus a ``CHEMACT`` line; ``FalloffRate::evalFromStruct``
  (``Falloff.h:154-176``) then returns k0 F / (1 + Pr).

Supported rate types: elementary Arrhenius (incl. ``negative-A``), three-body,
Lindemann, Troe, SRI and Tsang falloff, PLOG, Chebyshev, reaction ``orders``.
Anything else raises ``MechanismError`` naming the reaction.
"""
from __future__ import annotations

import math
import os
import re
import sys
from dataclasses import dataclass, field

import yaml

AVOGADRO = 6.02214076e26          # 1/kmol
BOLTZMANN = 1.380649e-23          # J/K
GAS_CONSTANT = AVOGADRO * BOLTZMANN
ONE_ATM = 101325.0
SMALL_NUMBER = 1e-300

# Elements.cpp:65-82 (the entries needed by the shipped mechanisms and common
# combustion sets)
ATOMIC_WEIGHTS = {
    "H": 1.008, "He": 4.002602, "C": 12.011, "N": 14.007, "O": 15.999,
    "F": 18.998403163, "Ne": 20.1797, "Si": 28.085, "S": 32.06,
    "Cl": 35.45, "Ar": 39.95, "Kr": 83.798, "Xe": 131.293, "E": 0.000545,
}

# Units.cpp knownUnits (subset: length, quantity, time, energy)
_LENGTH = {"m": 1.0, "cm": 0.01, "mm": 0.001}
_QUANTITY = {"kmol": 1.0, "mol": 1e-3, "molec": 1.0 / AVOGADRO}
_TIME = {"s": 1.0, "ms": 0.001, "min": 60.0, "hr": 3600.0}
_ENERGY = {"J": 1.0, "kJ": 1000.0, "cal": 4.184, "kcal": 4184.0, "eV": 1.602176634e-19 * 1.0}
_PRESSURE = {"Pa": 1.0, "kPa": 1e3, "MPa": 1e6, "bar": 1e5, "atm": ONE_ATM, "torr": ONE_ATM / 760.0,
             "dyn/cm^2": 0.1}


class MechanismError(ValueError):
    pass


def _num(v) -> float:
    if isinstance(v, bool):
        raise MechanismError(f"expected a number, got {v!r}")
    if isinstance(v, (int, float)):
        return float(v)
    return float(str(v).strip())


def _split_value_units(v):
    """'0.0 cal/mol' -> (0.0, 'cal/mol'); 12.5 -> (12.5, '')."""
    if isinstance(v, str):
        s = v.strip()
        if " " in s:
            a, b = s.split(" ", 1)
            return float(a), b.strip()
        return float(s), ""
    return _num(v), ""


def _energy_units_factor(u: str) -> tuple[float, str]:
    """Return (factor to J/kmol or to K, kind) for an activation-energy unit string.

    Mirrors Units(string) factor accumulation (``Units.cpp:134-192``): the
    factor starts at 1.0 and is multiplied by pow(unit_factor, exponent) for
    each group in order.
    """
    u = u.strip()
    if u == "K":
        return 1.0, "K"
    factor = 1.0
    start = 0
    groups = re.split(r"([*/])", u)
    sign = 1.0
    kind = None
    for g in groups:
        g = g.strip()
        if g == "/":
            sign = -1.0
            continue
        if g == "*":
            sign = 1.0
            continue
        if not g:
            continue
        expo = 1.0
        if "^" in g:
            g, e = g.split("^")
            expo = float(e)
        expo *= sign
        if g in _ENERGY:
            factor *= _ENERGY[g] ** expo
            kind = "energy"
        elif g in _QUANTITY:
            factor *= _QUANTITY[g] ** expo
        else:
            raise MechanismError(f"unsupported activation-energy unit {u!r}")
    del start
    if kind == "energy" and u.startswith("eV"):
        factor *= AVOGADRO
    return factor, "J/kmol"


@dataclass
class UnitSystem:
    """Default units of a YAML file or of one reaction (``UnitSystem``,
    ``Units.cpp:431-522``): factors to SI for bare numbers."""
    length: float = 1.0
    quantity: float = 1.0
    time: float = 1.0
    energy: float = 1.0
    pressure: float = 1.0
    ea_factor: float = 1.0      # multiplies an Ea value to J/kmol (or R for K)
    explicit_ea: bool = False   # activation-energy given, not derived from energy/quantity

    @classmethod
    def from_node(cls, node: dict | None) -> "UnitSystem":
        return cls().merged(node or {}, force=True)

    def merged(self, node: dict | None, force: bool = False) -> "UnitSystem":
        """This system with the overrides of a ``units`` map applied
        (``UnitSystem::setDefaults(map)``, ``Units.cpp:475-522``): a reaction's
        own ``units`` entry applies to that reaction's values only.  Without an
        explicit ``activation-energy``, Ea takes energy / quantity."""
        if not node and not force:
            return self
        node = node or {}
        us = UnitSystem(self.length, self.quantity, self.time, self.energy, self.pressure, self.ea_factor,
                        self.explicit_ea)
        for key in node:
            if key not in ("length", "quantity", "time", "energy", "pressure", "activation-energy", "mass",
                           "temperature", "current"):
                raise MechanismError(f"unsupported units entry {key!r}")
        try:
            if "length" in node:
                us.length = _LENGTH[node["length"]]
            if "quantity" in node:
                us.quantity = _QUANTITY[node["quantity"]]
            if "time" in node:
                us.time = _TIME[node["time"]]
            if "energy" in node:
                us.energy = _ENERGY[node["energy"]]
            if "pressure" in node:
                us.pressure = _PRESSURE[node["pressure"]]
        except KeyError as e:
            raise MechanismError(f"unsupported unit {e.args[0]!r}") from None
        if node.get("temperature", "K") != "K":
            raise MechanismError("temperature scales other than K are not supported")
        if "mass" in node and node["mass"] != "kg":
            raise MechanismError("mass units other than kg are not supported")
        if "activation-energy" in node:
            ea = node["activation-energy"]
            f, kind = _energy_units_factor(ea)
            us.ea_factor = GAS_CONSTANT if ea == "K" else f
            us.explicit_ea = True
        elif not us.explicit_ea:
            # m_activation_energy_factor = m_energy_factor / m_quantity_factor
            us.ea_factor = us.energy / us.quantity
        return us

    def rate_coeff(self, value, order: float) -> float:
        """A in the input system -> SI kmol/m^3/s for a reaction of total order n.

        ``UnitSystem::convertTo`` (``Units.cpp:562-571``) with destination
        units (kmol/m^3)^(1-n)/s: length dim 3(n-1), quantity dim 1-n, time -1.
        A value with explicit units (``2.9e-17 cm^3/molec/s``) is converted by
        those units alone (``Units(string)``, ``Units.cpp:134-192``).
        """
        v, u = _split_value_units(value)
        if u:
            return v * _rate_units_factor(u)
        L = 3.0 * (order - 1.0)
        Q = 1.0 - order
        Tm = -1.0
        return (v / 1.0 * (1.0 ** 0.0) * (self.length ** L) * (self.time ** Tm)
                * (self.quantity ** Q) * (1.0 ** 0.0) * (1.0 ** 0.0))

    def ea_over_r(self, value) -> float:
        v, u = _split_value_units(value)
        if u:
            f, kind = _energy_units_factor(u)
            if u == "K":
                return v
            return v * f / (GAS_CONSTANT * 1.0)
        return v * self.ea_factor / GAS_CONSTANT


def _rate_units_factor(u: str) -> float:
    """SI factor of a rate-coefficient unit string such as ``cm^3/molec/s``."""
    factor = 1.0
    sign = 1.0
    for g in re.split(r"([*/])", u.strip()):
        g = g.strip()
        if g in ("*", "/"):
            sign = -1.0 if g == "/" else 1.0
            continue
        if not g:
            continue
        expo = 1.0
        if "^" in g:
            g, e = g.split("^")
            expo = float(e)
        table = _LENGTH if g in _LENGTH else _QUANTITY if g in _QUANTITY else _TIME if g in _TIME else None
        if table is None:
            raise MechanismError(f"unsupported rate-coefficient unit {u!r}")
        factor *= table[g] ** (expo * sign)
    return factor


def _pressure(value, units: "UnitSystem | None" = None) -> float:
    v, u = _split_value_units(value)
    if not u:
        return v * (units.pressure if units is not None else 1.0)
    if u not in _PRESSURE:
        raise MechanismError(f"unsupported pressure unit {u!r}")
    return v * _PRESSURE[u]


@dataclass
class Species:
    name: str
    composition: dict
    mw: float
    tmin: float
    tmid: float
    tmax: float
    pref: float
    low: list
    high: list


@dataclass
class Reaction:
    equation: str
    rtype: int                 # 0 elementary, 1 three-body, 2 Lindemann, 3 Troe, 4 PLOG, 5 Chebyshev, 6 SRI, 7 Tsang
    reversible: bool
    A: float
    b: float
    EaR: float
    A0: float = 0.0
    b0: float = 0.0
    EaR0: float = 0.0
    troe: tuple = (0.0, 0.0, 0.0, 0.0)   # (a, 1/T3, 1/T1, T2)
    reac: list = field(default_factory=list)   # [(k, stoich, order)] in name order (+ non-reactant orders)
    prod: list = field(default_factory=list)   # [(k, stoich)] in name order
    tb_default: float = 0.0
    tb_eff: list = field(default_factory=list)  # [(k, eff - default)] sorted by k
    dn: float = 0.0
    plog: list = field(default_factory=list)    # PLOG: [(P [Pa], A, b, Ea/R)] sorted by P
    cheb: tuple = ()                            # Chebyshev: (Tmin, Tmax, Pmin, Pmax, rows)
    sri: tuple = ()                             # SRI: (a, b, c, d, e)
    tsang: tuple = ()                           # Tsang: (A, B)
    chem_act: bool = False                      # chemically activated falloff


TYPE_ELEMENTARY, TYPE_THREE_BODY, TYPE_LINDEMANN, TYPE_TROE, TYPE_PLOG, TYPE_CHEBYSHEV = 0, 1, 2, 3, 4, 5
TYPE_SRI, TYPE_TSANG = 6, 7
FALLOFF_TYPES = (TYPE_LINDEMANN, TYPE_TROE, TYPE_SRI, TYPE_TSANG)


def _parse_equation(eq: str):
    """parseReactionEquation (Reaction.cpp:889-955)."""
    tokens = eq.split()
    tokens.append("+")
    reactants: dict[str, float] = {}
    products: dict[str, float] = {}
    last_used = None
    in_reactants = True
    reversible = None
    for i in range(1, len(tokens)):
        t = tokens[i]
        if t == "+" or t.startswith("(+") or t in ("<=>", "=", "=>"):
            species = tokens[i - 1]
            stoich = 1.0
            if last_used is not None and tokens[last_used] == "(+" and species.endswith(")"):
                species = "(+" + species
            elif last_used == i - 1 and species.startswith("(+") and species.endswith(")"):
                pass
            elif (last_used is None and i == 1) or last_used == i - 2:
                pass
            elif (last_used is None and i == 2) or last_used == i - 3:
                stoich = float(tokens[i - 2])
            else:
                raise MechanismError(f"cannot parse reaction equation {eq!r}")
            target = reactants if in_reactants else products
            target[species] = target.get(species, 0.0) + stoich
            last_used = i
        if t in ("<=>", "="):
            reversible = True
            in_reactants = False
        elif t == "=>":
            reversible = False
            in_reactants = False
    if reversible is None:
        raise MechanismError(f"reaction {eq!r} has no '<=>', '=' or '=>'")
    return reactants, products, reversible


def _detect_third_body(reactants, products, rate_type, node):
    """Reaction::setEquation third-body detection (Reaction.cpp:389-555).

    Returns (third_body_name or None, mass_action) and edits the maps in place.
    """
    if rate_type == "elementary":
        return None, True
    third_body = None
    count = countM = countF = 0
    for name in sorted(reactants):
        stoich = reactants[name]
        if name in products:
            if countF == 0:
                third_body = name
            generic = third_body in ("M", "(+M)", "(+ M)")
            count += 1
            countM += int(generic)
            if stoich > 1 and products.get(third_body, 0) > 1:
                count += 1
                countM += int(generic)
            if name.startswith("(+") and name.endswith(")"):
                countF += 1
    explicit_effs = None
    if count == 0:
        if rate_type.startswith("three-body"):
            raise MechanismError(f"three-body reaction without collider: {node['equation']!r}")
        return None, True
    elif countF == 1:
        pass
    elif countM > 1:
        raise MechanismError("multiple generic third body colliders 'M' are not supported")
    elif count > 1:
        if countM:
            third_body = "M"
        elif "efficiencies" in node:
            effs = node["efficiencies"]
            if len(effs) != 1 or next(iter(effs)) not in reactants:
                raise MechanismError(f"ambiguous third body in {node['equation']!r}")
            third_body = next(iter(effs))
            explicit_effs = True
        else:
            if rate_type.startswith("three-body") or "default-efficiency" in node:
                raise MechanismError(f"ambiguous third body in {node['equation']!r}")
            return None, True
    elif third_body != "M" and not rate_type.startswith("three-body") and not third_body.startswith("(+"):
        nreac = nprod = 0
        for s in reactants.values():
            if math.trunc(s) != s:
                return None, True
            nreac += int(s)
        for s in products.values():
            if math.trunc(s) != s:
                return None, True
            nprod += int(s)
        if nreac != 3 and nprod != 3:
            return None, True
    del explicit_effs
    # adjust coefficients
    for side in (reactants, products):
        v = side[third_body]
        if math.trunc(v) != 1:
            side[third_body] = v - 1.0
        else:
            del side[third_body]
    name = third_body
    mass_action = True
    if name.startswith("(+ "):
        mass_action = False
        name = name[3:-1]
    elif name.startswith("(+"):
        mass_action = False
        name = name[2:-1]
    return name, mass_action


class _Yaml12Loader(yaml.SafeLoader):
    """SafeLoader with YAML 1.2 core-schema scalars, as yaml-cpp reads them.

    PyYAML follows YAML 1.1, where ``NO``/``yes``/``on`` are booleans and
    ``1e5`` is a string; species names such as ``NO`` must stay strings.
    """


_Yaml12Loader.yaml_implicit_resolvers = {
    k: [(tag, rx) for tag, rx in v if tag not in ("tag:yaml.org,2002:bool", "tag:yaml.org,2002:float")]
    for k, v in yaml.SafeLoader.yaml_implicit_resolvers.items()
}
_Yaml12Loader.add_implicit_resolver(
    "tag:yaml.org,2002:bool", re.compile(r"^(?:true|True|TRUE|false|False|FALSE)$"), list("tTfF"))
_Yaml12Loader.add_implicit_resolver(
    "tag:yaml.org,2002:float",
    re.compile(r"^(?:[-+]?(?:\.[0-9]+|[0-9]+(?:\.[0-9]*)?)(?:[eE][-+]?[0-9]+)?|[-+]?\.(?:inf|Inf|INF)|\.(?:nan|NaN|NAN))$"),
    list("-+0123456789."))


def _load_yaml(path):
    with open(path) as f:
        return yaml.load(f, Loader=_Yaml12Loader)


def _find_phase(doc, phase_name):
    phases = doc["phases"]
    if phase_name is None:
        return phases[0]
    for p in phases:
        if p["name"] == phase_name:
            return p
    raise MechanismError(f"phase {phase_name!r} not found")


def _collect_species(doc, phase, path, search_dirs=()):
    """Species list in phase order, resolving 'file.yaml/species' references."""
    base_dir = os.path.dirname(os.path.abspath(path))
    dirs = [base_dir] + list(search_dirs) + os.environ.get("CANTERA_DATA", "").split(os.pathsep)
    spec_entry = phase.get("species", "all")
    all_local = {s["name"]: s for s in doc.get("species", [])}
    out = []
    if spec_entry == "all":
        return list(doc["species"])
    if isinstance(spec_entry, list) and spec_entry and isinstance(spec_entry[0], str):
        return [all_local[n] for n in spec_entry]
    for item in spec_entry:
        for key, names in item.items():
            if key == "species":
                src = all_local
            else:
                fname, section = key.split("/", 1)
                cands = [os.path.join(d, fname) for d in dirs if d]
                found = [c for c in cands if os.path.exists(c)]
                if not found:
                    raise MechanismError(f"species source {fname!r} not found in {dirs}")
                other = _load_yaml(found[0])
                src = {s["name"]: s for s in other[section]}
            if names == "all":
                out.extend(src.values())
            else:
                out.extend(src[n] for n in names)
    return out


def compile_mechanism(path: str, phase_name: str | None = None, search_dirs=()):
    return compile_document(_load_yaml(path), path, phase_name, search_dirs)


def compile_document(doc: dict, path: str, phase_name: str | None = None, search_dirs=()):
    """Compile a mechanism already held as a Cantera-YAML document (dicts and
    lists); `path` names its source (for species references and messages)."""
    phase = _find_phase(doc, phase_name)
    if phase.get("thermo") != "ideal-gas":
        raise MechanismError(f"phase {phase['name']!r} is {phase.get('thermo')!r}; only ideal-gas is supported")
    units = UnitSystem.from_node(doc.get("units"))
    elements = list(phase["elements"])
    spec_nodes = _collect_species(doc, phase, path, search_dirs)
    species = []
    index = {}
    for s in spec_nodes:
        comp = {k: _num(v) for k, v in s["composition"].items()}
        wt = 0.0
        for e in elements:
            wt += comp.get(e, 0.0) * ATOMIC_WEIGHTS[e]
        for e in comp:
            if e not in elements:
                raise MechanismError(f"species {s['name']} uses element {e} not in phase")
        wt = max(wt, 1e-20)
        th = s["thermo"]
        if th.get("model") != "NASA7":
            raise MechanismError(f"species {s['name']}: thermo model {th.get('model')!r} not supported")
        tr = [float(x) for x in th["temperature-ranges"]]
        data = [[_num(x) for x in row] for row in th["data"]]
        pref = ONE_ATM
        if "reference-pressure" in th:
            pref = _num(th["reference-pressure"])
        if len(tr) == 3:
            low, high = data[0], data[1]
        else:
            low = high = data[0]
        index[s["name"]] = len(species)
        species.append(Species(s["name"], comp, wt, tr[0], tr[1], tr[-1], pref, low, high))

    rxn_nodes = []
    rsec = phase.get("reactions", "all")
    if rsec == "all" or rsec is None:
        rxn_nodes = list(doc.get("reactions", []))
    elif isinstance(rsec, list):
        for item in rsec:
            if isinstance(item, str):
                rxn_nodes.extend(doc[item])
            else:
                raise MechanismError("reaction section references to other files are not supported")
    elif rsec == "none":
        rxn_nodes = []
    else:
        rxn_nodes = list(doc.get(rsec, []))

    reactions = []
    for node in rxn_nodes:
        reactions.append(_compile_reaction(node, index, units))
    return {
        "phase": phase["name"], "elements": elements,
        "species": species, "reactions": reactions, "source": os.path.basename(path),
    }


def _compile_reaction(node, index, units: UnitSystem) -> Reaction:
    eq = node["equation"]
    # a reaction's own `units` map overrides the file defaults for every value
    # of this reaction (AnyMap::applyUnits on the reaction node)
    units = units.merged(node.get("units"))
    rate_type = node.get("type", "")
    if rate_type in ("", "reaction"):
        rate_type = ""
    chem_act = rate_type == "chemically-activated"
    if chem_act:
        rate_type = "falloff"
    if rate_type not in ("", "elementary", "three-body", "falloff", "pressure-dependent-Arrhenius", "Chebyshev"):
        raise MechanismError(f"reaction {eq!r}: rate type {rate_type!r} not supported")
    pdep = rate_type in ("pressure-dependent-Arrhenius", "Chebyshev")
    reactants, products, reversible = _parse_equation(eq)
    if pdep:
        # a (+M) written around a Chebyshev equation is the deprecated form;
        # PLOG and Chebyshev reactions have no third body
        for side in (reactants, products):
            for name in [n for n in side if n == "M" or n.startswith("(+")]:
                raise MechanismError(f"reaction {eq!r}: {rate_type} reactions take no third body")
        tb_name, mass_action = None, True
    else:
        tb_name, mass_action = _detect_third_body(reactants, products, rate_type, node)
    if rate_type == "falloff":
        if tb_name is None:
            raise MechanismError(f"falloff reaction {eq!r} has no (+M)")
        mass_action = False
    for name in list(reactants) + list(products):
        if name not in index:
            raise MechanismError(f"reaction {eq!r}: undeclared species {name!r}")
    orders = {k: _num(v) for k, v in (node.get("orders") or {}).items()}
    # Reaction::check (Reaction.cpp:135-162): explicit orders need their flags
    if not node.get("nonreactant-orders", False):
        for name in orders:
            if name not in reactants:
                raise MechanismError(f"reaction {eq!r}: order specified for non-reactant species {name!r}")
    if not node.get("negative-orders", False):
        for name, o in orders.items():
            if o < 0.0:
                raise MechanismError(f"reaction {eq!r}: negative order specified for species {name!r}")
    if orders and reversible:
        raise MechanismError(f"reaction {eq!r}: orders only allowed on irreversible reactions")
    # rate units: reactant orders (Reaction::calculateRateCoeffUnits)
    order = 0.0
    for name, o in orders.items():
        order += o
    for name, s in sorted(reactants.items()):
        if name not in orders:
            order += s
    is_tb = tb_name is not None
    if is_tb and mass_action:
        order += 1.0
    negA = bool(node.get("negative-A", False))

    def arr(key, extra_order):
        rc = node[key]
        if isinstance(rc, dict):
            A = units.rate_coeff(rc["A"], order + extra_order)
            b = _num(rc["b"])
            EaR = units.ea_over_r(rc["Ea"]) if "Ea" in rc else 0.0
        else:
            A = units.rate_coeff(rc[0], order + extra_order)
            b = _num(rc[1])
            EaR = units.ea_over_r(rc[2]) if len(rc) > 2 else 0.0
        if A < 0 and not negA:
            raise MechanismError(f"reaction {eq!r}: negative A requires 'negative-A: true'")
        return A, b, EaR

    r = Reaction(eq, TYPE_ELEMENTARY, reversible, 0.0, 0.0, 0.0)
    r.chem_act = chem_act
    if rate_type == "falloff":
        # FalloffRate::setParameters (Falloff.cpp:130-160): the low-pressure
        # rate is one concentration order above the high-pressure one; the
        # chemically activated form keeps the reaction's order on the low side
        lo = 0.0 if chem_act else 1.0
        r.A, r.b, r.EaR = arr("high-P-rate-constant", lo - 1.0)
        r.A0, r.b0, r.EaR0 = arr("low-P-rate-constant", lo)
        if "Troe" in node:
            t = node["Troe"]
            a = _num(t["A"])
            T3 = _num(t["T3"])
            T1 = _num(t["T1"])
            rt3 = math.inf if abs(T3) < SMALL_NUMBER else 1.0 / T3
            rt1 = math.inf if abs(T1) < SMALL_NUMBER else 1.0 / T1
            T2 = _num(t["T2"]) if "T2" in t else 0.0
            r.troe = (a, rt3, rt1, T2)
            r.rtype = TYPE_TROE
        elif "SRI" in node:
            # SriRate::setParameters / setFalloffCoeffs (Falloff.cpp:347-375, 405-428)
            f = node["SRI"]
            c = [_num(f["A"]), _num(f["B"]), _num(f["C"])]
            if "D" in f:
                c.append(_num(f["D"]))
            if "E" in f:
                c.append(_num(f["E"]))
            if len(c) not in (3, 5):
                raise MechanismError(f"reaction {eq!r}: SRI needs 3 or 5 coefficients, got {len(c)}")
            if c[2] < 0.0:
                raise MechanismError(f"reaction {eq!r}: SRI parameter C is less than zero")
            if len(c) == 5 and c[3] < 0.0:
                raise MechanismError(f"reaction {eq!r}: SRI parameter D is less than zero")
            if len(c) == 3:
                c += [1.0, 0.0]
            r.sri = tuple(c)
            r.rtype = TYPE_SRI
        elif "Tsang" in node:
            # TsangRate::setParameters / setFalloffCoeffs (Falloff.cpp:454-469, 496-512)
            f = node["Tsang"]
            r.tsang = (_num(f["A"]), _num(f["B"]) if "B" in f else 0.0)
            r.rtype = TYPE_TSANG
        else:
            r.rtype = TYPE_LINDEMANN
    elif rate_type == "pressure-dependent-Arrhenius":
        ents = []
        for rc in node["rate-constants"]:
            ents.append((_pressure(rc["P"], units), units.rate_coeff(rc["A"], order), _num(rc["b"]),
                         units.ea_over_r(rc["Ea"]) if "Ea" in rc else 0.0))
        if not ents:
            raise MechanismError(f"reaction {eq!r}: PLOG without rate constants")
        r.plog = sorted(ents, key=lambda e: e[0])      # stable: equal pressures keep input order
        r.rtype = TYPE_PLOG
    elif rate_type == "Chebyshev":
        lu = units
        Tr = [_num(_split_value_units(t)[0]) for t in node["temperature-range"]]
        Pr = [_pressure(x, units) for x in node["pressure-range"]]
        rows = [[_num(x) for x in row] for row in node["data"]]
        if not rows or any(len(row) != len(rows[0]) for row in rows):
            raise MechanismError(f"reaction {eq!r}: inconsistent Chebyshev coefficient matrix")
        rows[0][0] += math.log10(lu.rate_coeff(1.0, order))
        r.cheb = (Tr[0], Tr[1], Pr[0], Pr[1], rows)
        r.rtype = TYPE_CHEBYSHEV
    else:
        r.A, r.b, r.EaR = arr("rate-constant", 0.0)
        r.rtype = TYPE_THREE_BODY if is_tb else TYPE_ELEMENTARY

    # Kinetics::addReaction: reactants in name order, orders override, non-reactant
    # orders appended with zero stoichiometry (in orders-map name order).
    reac = []
    for name in sorted(reactants):
        s = reactants[name]
        reac.append([index[name], s, s])
    for name in sorted(orders):
        k = index[name]
        for e in reac:
            if e[0] == k:
                e[2] = orders[name]
                break
        else:
            reac.append([k, 0.0, orders[name]])
    r.reac = [tuple(e) for e in reac]
    r.prod = [(index[name], products[name]) for name in sorted(products)]

    dn = 0.0
    for name in sorted(products):
        dn += products[name]
    for name in sorted(reactants):
        dn -= reactants[name]
    r.dn = dn

    if is_tb:
        if tb_name == "M":
            default = _num(node.get("default-efficiency", 1.0))
            effs = {k: _num(v) for k, v in (node.get("efficiencies") or {}).items()}
        else:
            default = 0.0
            effs = {tb_name: 1.0}
            if "default-efficiency" in node and _num(node["default-efficiency"]) != 0.0:
                raise MechanismError(f"reaction {eq!r}: explicit collider needs default efficiency 0")
        eff_k = {}
        for name, e in effs.items():
            if name not in index:
                raise MechanismError(f"reaction {eq!r}: efficiency for undeclared species {name!r}")
            eff_k[index[name]] = e
        r.tb_default = default
        r.tb_eff = [(k, eff_k[k] - default) for k in sorted(eff_k)]
    r.reversible = reversible
    return r


def _f(x: float) -> str:
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    return repr(float(x))


def write_mech(mech, out_path):
    lines = ["CGMECH 1", f"phase {mech['phase']}", f"source {mech['source']}"]
    els = mech["elements"]
    lines.append(f"elements {len(els)} " + " ".join(f"{e} {_f(ATOMIC_WEIGHTS[e])}" for e in els))
    lines.append(f"species {len(mech['species'])}")
    for i, s in enumerate(mech["species"]):
        comp = " ".join(_f(s.composition.get(e, 0.0)) for e in els)
        lines.append(" ".join([
            "S", str(i), s.name, _f(s.mw), _f(s.tmin), _f(s.tmid), _f(s.tmax), _f(s.pref),
            " ".join(_f(c) for c in s.low), " ".join(_f(c) for c in s.high), comp]))
    lines.append(f"reactions {len(mech['reactions'])}")
    for i, r in enumerate(mech["reactions"]):
        toks = ["R", str(i), str(r.rtype), str(int(r.reversible)), _f(r.dn),
                _f(r.A), _f(r.b), _f(r.EaR), _f(r.A0), _f(r.b0), _f(r.EaR0),
                " ".join(_f(x) for x in r.troe), str(len(r.reac))]
        for k, s, o in r.reac:
            toks += [str(k), _f(s), _f(o)]
        toks.append(str(len(r.prod)))
        for k, s in r.prod:
            toks += [str(k), _f(s)]
        toks += [_f(r.tb_default), str(len(r.tb_eff))]
        for k, e in r.tb_eff:
            toks += [str(k), _f(e)]
        lines.append(" ".join(toks))
        if r.chem_act:
            lines.append("CHEMACT")
        if r.rtype == TYPE_PLOG:
            lines.append(" ".join(["PLOG", str(len(r.plog))] + [_f(x) for e in r.plog for x in e]))
        elif r.rtype == TYPE_SRI:
            lines.append(" ".join(["SRI"] + [_f(x) for x in r.sri]))
        elif r.rtype == TYPE_TSANG:
            lines.append(" ".join(["TSANG"] + [_f(x) for x in r.tsang]))
        elif r.rtype == TYPE_CHEBYSHEV:
            Tmin, Tmax, Pmin, Pmax, rows = r.cheb
            lines.append(" ".join(["CHEB", _f(Tmin), _f(Tmax), _f(Pmin), _f(Pmax), str(len(rows)), str(len(rows[0]))]
                                  + [_f(x) for row in rows for x in row]))
        lines.append("E " + r.equation)
    lines.append("end")
    with open(out_path, "w") as f:
        f.write("\n".join(lines) + "\n")


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="compile a Cantera YAML mechanism to .mech tables")
    ap.add_argument("yaml")
    ap.add_argument("out")
    ap.add_argument("--phase", default=None)
    ap.add_argument("--search", action="append", default=[],
                    help="extra directory for 'file.yaml/species' references")
    a = ap.parse_args(argv)
    mech = compile_mechanism(a.yaml, a.phase, a.search)
    write_mech(mech, a.out)
    names = {TYPE_THREE_BODY: "three-body", TYPE_LINDEMANN: "Lindemann", TYPE_TROE: "Troe", TYPE_SRI: "SRI",
             TYPE_TSANG: "Tsang", TYPE_PLOG: "PLOG", TYPE_CHEBYSHEV: "Chebyshev"}
    counts = [f"{sum(r.rtype == t for r in mech['reactions'])} {n}" for t, n in names.items()
              if t in (TYPE_THREE_BODY, TYPE_LINDEMANN, TYPE_TROE) or any(r.rtype == t for r in mech["reactions"])]
    print(f"{a.out}: {len(mech['species'])} species, {len(mech['reactions'])} reactions ({', '.join(counts)})",
          file=sys.stderr)


if __name__ == "__main__":
    main()
