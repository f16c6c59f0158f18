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


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    return oracle
