"""Synthetic workloads for the benchmark configurations (seeded, reproducible).

Mixtures follow ``ThermoPhase::setEquivalenceRatio`` (src/thermo/ThermoPhase.cpp:879-910)
on a molar basis for a single-species fuel and the oxidizer O2:1, N2:3.76: the
oxygen demand of one mole of fuel is C + H/4 - O/2, and the fuel moles per mole
of O2 are phi divided by that demand.

Random numbers come from numpy's counter-based Philox generator with seed
0x5EED0000 + config number.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

ONE_ATM = 101325.0
SEED_BASE = 0x5EED0000


def rng(config: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(SEED_BASE + config))


class MechInfo:
    """Species names, molecular weights and element composition read from a
    compiled .mech file (host-only; no GPU library needed)."""

    def __init__(self, path: str):
        self.path = path
        self.species_names, mw, comp = [], [], []
        with open(path) as f:
            for line in f:
                t = line.split()
                if not t:
                    continue
                if t[0] == "elements":
                    E = int(t[1])
                    self.elements = [t[2 + 2 * i] for i in range(E)]
                elif t[0] == "S":
                    self.species_names.append(t[2])
                    mw.append(float(t[3]))
                    comp.append([float(x) for x in t[len(t) - len(self.elements):]])
        self.molecular_weights = np.array(mw)
        self.composition = np.array(comp)
        self.n_species = len(self.species_names)
        self._index = {n: i for i, n in enumerate(self.species_names)}

    def species_index(self, name: str) -> int:
        return self._index[name]

    def mass_fractions_from_moles(self, X) -> np.ndarray:
        Y = np.asarray(X, dtype=np.float64) * self.molecular_weights
        return Y / Y.sum(axis=-1, keepdims=True)


def _fuel_demand(sol, fuel: str) -> float:
    """Moles of O2 needed per mole of fuel (complete oxidation to CO2 and H2O)."""
    comp = _composition(sol, fuel)
    return comp.get("C", 0.0) + comp.get("H", 0.0) / 4.0 - comp.get("O", 0.0) / 2.0


def _composition(sol, species: str) -> dict:
    # parse the element composition from the compiled mechanism file
    with open(sol.path) as f:
        elements = None
        for line in f:
            t = line.split()
            if not t:
                continue
            if t[0] == "elements":
                E = int(t[1])
                elements = [t[2 + 2 * i] for i in range(E)]
            elif t[0] == "S" and t[2] == species:
                vals = [float(x) for x in t[-len(elements):]]
                return {e: v for e, v in zip(elements, vals) if v}
    raise KeyError(species)


def equivalence_ratio_moles(sol, fuel: str, phi, oxidizer=(("O2", 1.0), ("N2", 3.76))) -> np.ndarray:
    """Mole fractions (n, K) of fuel/air mixtures at equivalence ratios phi."""
    phi = np.atleast_1d(np.asarray(phi, dtype=np.float64))
    X = np.zeros((phi.shape[0], sol.n_species))
    demand = _fuel_demand(sol, fuel)
    o2 = dict(oxidizer).get("O2", dict(oxidizer).get("o2"))
    for name, v in oxidizer:
        X[:, sol.species_index(name)] += v
    X[:, sol.species_index(fuel)] += phi * o2 / demand
    return X / X.sum(axis=1, keepdims=True)


@dataclass
class ReactorWorkload:
    name: str
    T0: np.ndarray
    P: np.ndarray
    Y0: np.ndarray          # (n, K)
    dt: np.ndarray          # (n,)
    nsteps: int
    rtol: float
    atol: float


def config1(sol, n=1024) -> ReactorWorkload:
    """h2o2, 1024 reactors to 1e-3 s, outputs every 1e-5 s, rtol 1e-8."""
    g = rng(1)
    T0 = g.uniform(1000.0, 1800.0, n)
    phi = g.uniform(0.5, 2.0, n)
    X = equivalence_ratio_moles(sol, "H2", phi)
    Y = sol.mass_fractions_from_moles(X)
    return ReactorWorkload("h2o2 IGCPR, 1024 reactors, 100 x 1e-5 s", T0, np.full(n, ONE_ATM), Y,
                           np.full(n, 1e-5), 100, 1e-8, 1e-15)


def config2_states(sol, n=1_048_576, config=2):
    """gri30 random states: returns T (n,), P (n,), Y (K, n) species-major."""
    g = rng(config)
    K = sol.n_species
    T = g.uniform(300.0, 2500.0, n)
    P = ONE_ATM * np.exp(g.uniform(np.log(0.1), np.log(50.0), n))
    phi = g.uniform(0.3, 3.0, n)
    X = np.empty((n, K))
    chunk = 1 << 16
    majors = [sol.species_index(s) for s in ("CH4", "O2", "N2")]
    minor = np.ones(K, dtype=bool)
    minor[majors] = False
    nminor = int(minor.sum())
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        Xm = equivalence_ratio_moles(sol, "CH4", phi[a:b])
        Xm[:, minor] = 10.0 ** g.uniform(-12.0, -2.0, (b - a, nminor))
        Xm /= Xm.sum(axis=1, keepdims=True)
        X[a:b] = Xm
    Y = sol.mass_fractions_from_moles(X)
    return T, P, np.ascontiguousarray(Y.T)


def config3(sol, n=65536, nsteps=100, seed=None) -> ReactorWorkload:
    """gri30, constant-pressure reactors of uneven stiffness, per-reactor dt.
    `seed` (default: this workload's own) draws another set of reactors from
    the same distribution."""
    g = rng(3) if seed is None else np.random.Generator(np.random.Philox(seed))
    T0 = g.uniform(900.0, 2000.0, n)
    phi = g.uniform(0.5, 3.0, n)
    P = ONE_ATM * np.exp(g.uniform(np.log(1.0), np.log(40.0), n))
    dt = np.exp(g.uniform(np.log(1e-6), np.log(1e-4), n))
    X = equivalence_ratio_moles(sol, "CH4", phi)
    return ReactorWorkload(f"gri30 IGCPR, {n} reactors x {nsteps} steps, dt log-U[1e-6,1e-4] s",
                           T0, P, sol.mass_fractions_from_moles(X), dt, nsteps, 1e-8, 1e-15)


def config5(sol, n=262144, nsteps=100) -> ReactorWorkload:
    """nDodecane, constant-pressure reactors at 20-60 atm, dt = 1e-6 s."""
    g = rng(5)
    T0 = g.uniform(700.0, 2200.0, n)
    phi = g.uniform(0.3, 3.0, n)
    P = ONE_ATM * g.uniform(20.0, 60.0, n)
    X = equivalence_ratio_moles(sol, "c12h26", phi, oxidizer=(("o2", 1.0), ("n2", 3.76)))
    return ReactorWorkload(f"nDodecane IGCPR, {n} reactors x {nsteps} steps, dt 1e-6 s",
                           T0, P, sol.mass_fractions_from_moles(X), np.full(n, 1e-6), nsteps, 1e-8, 1e-15)
