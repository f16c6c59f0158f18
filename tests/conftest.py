import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

MECH_DIR = os.path.join(ROOT, "cantera_amd", "mechanisms")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


def mech_file(name):
    return os.path.join(MECH_DIR, name + ".mech")


def parse_stoich(path):
    """Net stoichiometric matrix nu (K x I) and |nu| from a .mech file."""
    K = I = 0
    rows = []
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            if t[0] == "species":
                K = int(t[1])
            elif t[0] == "reactions":
                I = int(t[1])
            elif t[0] == "R":
                i = int(t[1])
                p = 15
                nr = int(t[p]); p += 1
                reac = [(int(t[p + 3 * j]), float(t[p + 3 * j + 1])) for j in range(nr)]
                p += 3 * nr
                np_ = int(t[p]); p += 1
                prod = [(int(t[p + 2 * j]), float(t[p + 2 * j + 1])) for j in range(np_)]
                rows.append((i, reac, prod))
    nu_r = np.zeros((K, I))
    nu_p = np.zeros((K, I))
    for i, reac, prod in rows:
        for k, s in reac:
            nu_r[k, i] += s
        for k, s in prod:
            nu_p[k, i] += s
    return nu_r, nu_p


def parse_third_bodies(path):
    """[(default efficiency, [(k, efficiency - default)])] of the reactions with
    a third body, from a .mech file."""
    out = []
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t or t[0] != "R":
                continue
            p = 15
            nr = int(t[p]); p += 1 + 3 * nr
            np_ = int(t[p]); p += 1 + 2 * np_
            tbdef, ntb = float(t[p]), int(t[p + 1])
            eff = [(int(t[p + 2 + 2 * j]), float(t[p + 3 + 2 * j])) for j in range(ntb)]
            if tbdef != 0.0 or ntb:
                out.append((tbdef, eff))
    return out


def extend_efficiencies(src, dst, seed=7, base=9, span=14):
    """Write a copy of the .mech file src to dst in which every reaction with a
    third body lists base..base+span-1 explicit efficiencies (reaction j gets
    base + j % span): the species it already lists, then species not yet
    listed with (efficiency - default) drawn from [0.1, 2].  Longer lists than
    any shipped mechanism holds, for the kernels' loops over efficiencies.
    Returns {reaction index: (default, [(k, efficiency - default)])}."""
    g = np.random.default_rng(seed)
    K = 0
    out, lists = [], {}
    with open(src) as f:
        for line in f:
            t = line.split()
            if t and t[0] == "species":
                K = int(t[1])
            if not t or t[0] != "R":
                out.append(line)
                continue
            p = 15
            nr = int(t[p]); p += 1 + 3 * nr
            np_ = int(t[p]); p += 1 + 2 * np_
            tbdef, ntb = float(t[p]), int(t[p + 1])
            if tbdef == 0.0 and ntb == 0:
                out.append(line)
                continue
            j = int(t[1])
            eff = [(int(t[p + 2 + 2 * i]), float(t[p + 3 + 2 * i])) for i in range(ntb)]
            have = {k for k, _ in eff}
            for k in g.permutation(K):
                if len(eff) >= base + j % span:
                    break
                if int(k) not in have:
                    eff.append((int(k), float(g.uniform(0.1, 2.0))))
            rest = t[p + 2 + 2 * ntb:]
            toks = t[:p] + [repr(tbdef), str(len(eff))] + [x for k, e in eff for x in (str(k), repr(e))] + rest
            out.append(" ".join(toks) + "\n")
            lists[j] = (tbdef, eff)
    with open(dst, "w") as f:
        f.writelines(out)
    return lists


def third_body_condition(path, X):
    """Cancellation factor of [M] = d ctot + sum (e_k - d) C_k for mole
    fractions X: max over reactions of (d + sum |e_k - d| X_k) / |d + sum (e_k - d) X_k|.
    Rounding in [M] is amplified by this factor in every rate that uses it
    (the reference sums [M] the same way, ThirdBodyCalc.h:75-91)."""
    c = 1.0
    for d, eff in parse_third_bodies(path):
        num = d + sum(abs(e) * X[k] for k, e in eff)
        den = abs(d + sum(e * X[k] for k, e in eff))
        c = max(c, num / max(den, 1e-300))
    return c


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    return oracle
